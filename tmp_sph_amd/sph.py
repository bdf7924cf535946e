"""Host-side interface of the MI355X delta-SPH step.

Two layers, both thin wrappers over the C ABI (include/sph_capi.h):

* the reference call surface -- ``calcHash``, ``sortParticles``, ``reorder``,
  ``SPH_NS_simpleversion``, ``PC_prediction``, ``PC_correction`` and
  ``set_params`` -- on caller-owned device arrays in the reference's layouts
  (src/header/mesh.cuh:20-29, src/header/ns.cuh:4, src/header/sph.cuh:26-28);
* :class:`Simulation`, the fast path: sorted SoA state with FP64 positions,
  whole predictor-corrector steps, hipGraph replay.

Cases: :func:`dambreak_cfg` builds the lattice dam-break configurations and
:func:`reference_params` the reference's own parameter recipe.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DamBreakCfg, SphParams, SphTuning, call, load

__all__ = [
    "DeviceArray", "SphParams", "DamBreakCfg", "SphTuning", "Simulation", "dambreak_cfg", "make_case",
    "reference_params", "subdivide_params", "set_params", "get_params", "calcHash", "sortParticles", "reorder",
    "SPH_NS_simpleversion", "PC_prediction", "PC_correction", "device_count", "synchronize",
]


def device_count() -> int:
    c = C.c_int()
    call("sph_device_count", C.byref(c))
    return c.value


def synchronize() -> None:
    call("sph_device_synchronize")


class DeviceArray:
    """A device allocation with a dtype and shape, freed with the object."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("sph_malloc", C.byref(p), self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        call("sph_memcpy_htod", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        call("sph_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes)
        return out

    def fill_bytes(self, value: int) -> None:
        call("sph_memset", C.c_void_p(self.ptr), int(value), self.nbytes)

    @property
    def c(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)

    def __del__(self):
        try:
            if self.ptr and _lib._lib is not None:
                _lib._lib.sph_free(C.c_void_p(self.ptr))
        except Exception:
            pass
        self.ptr = None


def _p(a) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, DeviceArray):
        return a.c
    raise TypeError("device arrays are DeviceArray instances")


# ----------------------------------------------------------------- parameters
def set_params(p: SphParams) -> None:
    """Replaces ``cudaMemcpyToSymbol(par, ...)`` (src/sph/sph_run.cu:105)."""
    call("sph_set_params", C.byref(p))


def get_params() -> SphParams:
    p = SphParams()
    call("sph_get_params", C.byref(p))
    return p


def reference_params(dx: float = 0.01, num_particles: int = 1140638, num_fluid: int = 669735) -> SphParams:
    """The reference's parameter recipe (src/sph/sph_run.cu:60-101)."""
    p = SphParams()
    call("sph_params_reference", C.byref(p), C.c_float(dx), int(num_particles), int(num_fluid))
    return p


def dambreak_cfg(dim: int = 2, dx: float = 0.01, tank=(4.0, 1.0, 1.0), column=(1.0, 1.0, 0.5), *,
                 eos: int = _lib.SPH_EOS_TAIT, wall_layers: int = 3, h_over_dx: float = 1.3,
                 rho0: float = 1000.0, g: float = -9.81, c0: float = 0.0, cfl: float = 0.2,
                 delta: float = 0.1, alpha: float = 0.02, gamma: float = 7.0, grid_pad: float = 0.1,
                 cell_over_kh: float = 1.0) -> DamBreakCfg:
    """Dam-break case: water column at the x=0 (and y=0) corner of the tank.

    Defaults are the 2-D configuration (1.0 m x 0.5 m column in a 4.0 m x 1.0 m
    tank, h = 1.3 dx, Tait gamma = 7, c0 = 10 sqrt(2 g H), delta = 0.1,
    alpha = 0.02, CFL 0.2, 3 wall layers).  In 2-D the plane is x-z.
    """
    c = DamBreakCfg()
    c.dim, c.eos, c.wall_layers = dim, eos, wall_layers
    c.dx = dx
    for k in range(3):
        c.tank[k] = tank[k]
        c.column[k] = column[k]
    c.h_over_dx, c.rho0, c.g, c.c0, c.cfl = h_over_dx, rho0, g, c0, cfl
    c.delta, c.alpha, c.gamma, c.grid_pad, c.cell_over_kh = delta, alpha, gamma, grid_pad, cell_over_kh
    return c


def subdivide_params(p: SphParams, div: int) -> SphParams:
    """The parameters on the cell grid divided by ``div`` (1, 2 or 4): the grid
    a context searches with tuning ``cell_div = div``."""
    out = SphParams()
    call("sph_params_subdivide", C.byref(p), int(div), C.byref(out))
    return out


def case_counts(cfg: DamBreakCfg) -> tuple[int, int]:
    nt, nf = C.c_int(), C.c_int()
    call("sph_case_dambreak_count", C.byref(cfg), C.byref(nt), C.byref(nf))
    return nt.value, nf.value


@dataclass
class Case:
    params: SphParams
    pos: np.ndarray    # float32 (N, 3)
    vel: np.ndarray    # float32 (N, 3)
    rho: np.ndarray    # float32 (N,)
    p: np.ndarray      # float32 (N,)
    type: np.ndarray   # int32 (N,)
    n_fluid: int

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def copy(self) -> "Case":
        return Case(self.params.copy(), self.pos.copy(), self.vel.copy(), self.rho.copy(), self.p.copy(),
                    self.type.copy(), self.n_fluid)


def make_case(cfg: DamBreakCfg) -> Case:
    """Generates the lattice case on the host (walls first, then fluid)."""
    nt, nf = case_counts(cfg)
    pos = np.zeros((nt, 3), np.float32)
    vel = np.zeros((nt, 3), np.float32)
    rho = np.zeros(nt, np.float32)
    prs = np.zeros(nt, np.float32)
    typ = np.zeros(nt, np.int32)
    prm = SphParams()
    a, b = C.c_int(), C.c_int()
    call("sph_case_dambreak", C.byref(cfg), nt, pos.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p),
         rho.ctypes.data_as(C.c_void_p), prs.ctypes.data_as(C.c_void_p), typ.ctypes.data_as(C.c_void_p),
         C.byref(prm), C.byref(a), C.byref(b))
    return Case(prm, pos, vel, rho, prs, typ, nf)


# ------------------------------------------------- reference call surface
def calcHash(pos: DeviceArray, particleHash: DeviceArray, particleIndex: DeviceArray, numParticles: int) -> None:
    """src/lib/mesh.cu:56-70."""
    call("sph_calcHash", _p(pos), _p(particleHash), _p(particleIndex), int(numParticles))


def sortParticles(particleHash: DeviceArray, particleIndex: DeviceArray, numParticles: int) -> None:
    """src/lib/mesh.cu:73-86 (stable ascending key sort carrying the indices)."""
    call("sph_sortParticles", _p(particleHash), _p(particleIndex), int(numParticles))


def reorder(oldPos, oldVel, olddensity, oldpressure, oldparticle_type, oldparticle_zone, sortedPos, sortedVel,
            sorteddensity, sortedpressure, sortedparticle_type, sortedparticle_zone, particleHash, particleIndex,
            dcellStart, dCellEnd, numParticles, table_l=None, table_r=None) -> None:
    """opt_src_v0/lib/mesh.cu:161-177 (cell ranges + gather + permutation table)."""
    call("sph_reorder", _p(oldPos), _p(oldVel), _p(olddensity), _p(oldpressure), _p(oldparticle_type),
         _p(oldparticle_zone), _p(sortedPos), _p(sortedVel), _p(sorteddensity), _p(sortedpressure),
         _p(sortedparticle_type), _p(sortedparticle_zone), _p(particleHash), _p(particleIndex), _p(dcellStart),
         _p(dCellEnd), int(numParticles), _p(table_l), _p(table_r))


def recover(pos, vel, density, pressure, numParticles, particleIndex, particle_type, particle_zone, sortedpos,
            sortedvel, sorteddensity, sortedpressure, sorted_particle_type, sorted_particle_zone) -> None:
    """src/lib/mesh.cu:181-194 (scatter sorted rows back to original order; src variant)."""
    call("sph_recover", _p(pos), _p(vel), _p(density), _p(pressure), int(numParticles), _p(particleIndex),
         _p(particle_type), _p(particle_zone), _p(sortedpos), _p(sortedvel), _p(sorteddensity), _p(sortedpressure),
         _p(sorted_particle_type), _p(sorted_particle_zone))


def SPH_NS_simpleversion(sortedPos, sortedVel, sorteddensity, sortedpressure, sorted_particle_type, densitydt,
                         Veldt, cellStart, cellEnd, numParticles, particleHash, timestep, dofv, rhop_sum=None,
                         w_sum=None) -> None:
    """opt_src_v1/equation/ns.cu:5-28: kernel A then kernel B."""
    call("sph_SPH_NS_simpleversion", _p(sortedPos), _p(sortedVel), _p(sorteddensity), _p(sortedpressure),
         _p(sorted_particle_type), _p(densitydt), _p(Veldt), _p(cellStart), _p(cellEnd), int(numParticles),
         _p(particleHash), int(timestep), _p(dofv), _p(rhop_sum), _p(w_sum))


def set_ns_mode(mode: int) -> None:
    """Kernels behind SPH_NS_simpleversion: 0 the row-staged fast loops
    (default), 1 the reference's FP32 operation order."""
    call("sph_set_ns_mode", int(mode))


def reserve_scratch(n: int) -> None:
    """Allocates the scratch of sortParticles and SPH_NS_simpleversion for n
    particles now (after set_params), so those calls do not allocate."""
    call("sph_reserve_scratch", int(n))


def get_ns_mode() -> int:
    m = C.c_int()
    call("sph_get_ns_mode", C.byref(m))
    return m.value


def PC_prediction(sortedPos, sortedVel, sorteddensity, sortedpressure, densitydt, Veldt, particleHash,
                  particleIndex, Pos_tmp, Vel_tmp, density_tmp, numParticles, sorted_particle_type) -> None:
    """opt_src_v0/sph/sph_run.cu:486-500."""
    call("sph_PC_prediction", _p(sortedPos), _p(sortedVel), _p(sorteddensity), _p(sortedpressure), _p(densitydt),
         _p(Veldt), _p(particleHash), _p(particleIndex), _p(Pos_tmp), _p(Vel_tmp), _p(density_tmp),
         int(numParticles), _p(sorted_particle_type))


def PC_correction(sortedPos, sortedVel, sorteddensity, sortedpressure, densitydt, Veldt, particleHash,
                  particleIndex, Pos_tmp, Vel_tmp, density_tmp, numParticles, sorted_particle_type) -> None:
    """opt_src_v0/sph/sph_run.cu:502-514."""
    call("sph_PC_correction", _p(sortedPos), _p(sortedVel), _p(sorteddensity), _p(sortedpressure), _p(densitydt),
         _p(Veldt), _p(particleHash), _p(particleIndex), _p(Pos_tmp), _p(Vel_tmp), _p(density_tmp),
         int(numParticles), _p(sorted_particle_type))


# ---------------------------------------------------------------- fast path
def _hp(a: np.ndarray | None) -> C.c_void_p:
    return C.c_void_p(0) if a is None else a.ctypes.data_as(C.c_void_p)


class Simulation:
    """Dam-break simulation(s) on the current device (sorted SoA state, FP64 positions).

    ``Simulation(case)`` runs one case; ``Simulation([case0, case1, ...])``
    runs the cases as one batch (one launch per phase for all of them), each
    with its own parameters and time step.  Arrays returned by :meth:`state`
    hold the cases back to back; :attr:`offsets` gives their start indices.
    """

    def __init__(self, cases: "Case | list[Case]"):
        load()
        batch = list(cases) if isinstance(cases, (list, tuple)) else [cases]
        if not batch:
            raise ValueError("no cases")
        self.cases_params = [c.params.copy() for c in batch]
        self.params = self.cases_params[0]
        counts = np.array([c.n for c in batch], np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.n = int(counts.sum())
        self.n_fluid = int(sum(c.n_fluid for c in batch))
        self.K = len(batch)
        cat = lambda f, d: np.ascontiguousarray(np.concatenate([getattr(c, f) for c in batch]), d)  # noqa: E731
        prm = (SphParams * self.K)(*self.cases_params)
        h = C.c_void_p()
        call("sph_sim_create_batch", prm, self.K, counts.ctypes.data_as(C.POINTER(C.c_int)),
             _hp(cat("pos", np.float32)), _hp(cat("vel", np.float32)), _hp(cat("rho", np.float32)),
             _hp(cat("p", np.float32)), _hp(cat("type", np.int32)), C.byref(h))
        self._h = h

    @classmethod
    def restore(cls, path) -> "Simulation":
        """A context continuing bit-identically from a :meth:`checkpoint` file."""
        load()
        h = C.c_void_p()
        call("sph_sim_restore", os.fsencode(os.fspath(path)), C.byref(h))
        self = cls.__new__(cls)
        self._h = h
        K = C.c_int()
        call("sph_sim_num_instances", h, C.byref(K))
        self.K = K.value
        self.cases_params = []
        for k in range(self.K):
            p = SphParams()
            call("sph_sim_instance_params", h, k, C.byref(p))
            self.cases_params.append(p)
        self.params = self.cases_params[0]
        counts = np.array([p.numParticles for p in self.cases_params], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.n = int(counts.sum())
        self.n_fluid = int(np.count_nonzero(self.state()["type"] == 1))
        return self

    def checkpoint(self, path) -> None:
        """Writes the complete state (sorted order, FP64 positions, step count)."""
        call("sph_sim_checkpoint", self._h, os.fsencode(os.fspath(path)))

    def output_open(self, directory, binary: bool = True, rho: bool = False, type: bool = False) -> None:
        """Starts the asynchronous frame writer (files directory/fluid_t=<frame>.vtk)."""
        os.makedirs(directory, exist_ok=True)
        call("sph_sim_output_open", self._h, os.fsencode(os.fspath(directory)), int(bool(binary)),
             int(bool(rho)) | (int(bool(type)) << 1))

    def output_frame(self, frame: int) -> None:
        """Queues the current state as frame ``frame``; returns before it is written."""
        call("sph_sim_output_frame", self._h, int(frame))

    def output_flush(self) -> int:
        """Waits for every queued frame; returns the number written so far."""
        n = C.c_longlong()
        call("sph_sim_output_flush", self._h, C.byref(n))
        return n.value

    def output_close(self) -> None:
        call("sph_sim_output_close", self._h)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib._lib.sph_sim_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def tuning(self) -> SphTuning:
        t = SphTuning()
        call("sph_sim_get_tuning", self._h, C.byref(t))
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning
        for k, v in kw.items():
            setattr(t, k, v)
        call("sph_sim_set_tuning", self._h, C.byref(t))

    def step(self, n: int = 1) -> None:
        call("sph_sim_step", self._h, int(n))

    def capture(self) -> None:
        call("sph_sim_capture", self._h)

    def replay(self, n: int = 1) -> None:
        call("sph_sim_replay", self._h, int(n))

    def synchronize(self) -> None:
        call("sph_sim_synchronize", self._h)

    def state(self) -> dict:
        n = self.n
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        rho = np.empty(n, np.float32)
        prs = np.empty(n, np.float32)
        typ = np.empty(n, np.int32)
        call("sph_sim_download", self._h, _hp(pos), _hp(vel), _hp(rho), _hp(prs), _hp(typ))
        return {"pos": pos, "vel": vel, "rho": rho, "p": prs, "type": typ}

    def pos64(self) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        call("sph_sim_download_pos64", self._h, _hp(out))
        return out

    def probe(self, band: float) -> tuple[float, float]:
        """Surge front (max x of in-domain fluid) and column height (max z of
        fluid with x <= band); the maxima over all instances of a batch."""
        f, h = C.c_double(), C.c_double()
        call("sph_sim_probe", self._h, C.c_double(band), C.byref(f), C.byref(h))
        return f.value, h.value

    def probe_batch(self, band: float) -> tuple[np.ndarray, np.ndarray]:
        f = np.zeros(self.K)
        h = np.zeros(self.K)
        call("sph_sim_probe_batch", self._h, C.c_double(band), _hp(f), _hp(h))
        return f, h

    def probe_history(self, k: int, intervals: int, band: float) -> tuple[np.ndarray, np.ndarray]:
        """Advances `intervals` x k steps (graph replays once a step is
        captured), probing front and height after every interval on the
        device; the host waits once.  Returns (front, height), each
        [intervals] for one instance or [intervals, K] for a batch."""
        f = np.zeros((max(intervals, 1), self.K))
        h = np.zeros((max(intervals, 1), self.K))
        call("sph_sim_probe_history", self._h, int(k), int(intervals), C.c_double(band), _hp(f), _hp(h))
        f, h = f[:intervals], h[:intervals]
        return (f[:, 0], h[:, 0]) if self.K == 1 else (f, h)

    def autotune(self, reps: int = 3) -> tuple[dict, float]:
        """Picks the fastest neighbour-kernel launch shape for the current state."""
        t = SphTuning()
        us = C.c_double()
        call("sph_sim_autotune", self._h, int(reps), C.byref(t), C.byref(us))
        return t.as_dict(), us.value

    def eval_rates(self) -> dict:
        n = self.n
        out = {"drho": np.empty(n, np.float32), "acc": np.empty((n, 3), np.float32),
               "dofv": np.empty(n, np.float32), "p": np.empty(n, np.float32), "rho": np.empty(n, np.float32),
               "keys": np.empty(n, np.int32), "perm": np.empty(n, np.int32)}
        call("sph_sim_eval", self._h, _hp(out["drho"]), _hp(out["acc"]), _hp(out["dofv"]), _hp(out["p"]),
             _hp(out["rho"]), _hp(out["keys"]), _hp(out["perm"]))
        return out

    def gauges(self, gx, gy, zcap, halfw: float) -> np.ndarray:
        """Height gauges (largest fluid z in each band, instance 0)."""
        gx = np.ascontiguousarray(gx, np.float32)
        gy = np.ascontiguousarray(gy, np.float32)
        zc = np.ascontiguousarray(zcap, np.float32)
        out = np.empty(gx.size, np.float32)
        call("sph_sim_gauges", self._h, int(gx.size), _hp(gx), _hp(gy), _hp(zc), C.c_float(halfw), _hp(out))
        return out

    def pressure_at(self, ids) -> np.ndarray:
        """Pressure of the particles with the given original indices."""
        ids = np.ascontiguousarray(ids, np.int32)
        out = np.empty(ids.size, np.float32)
        call("sph_sim_pressure_at", self._h, int(ids.size), _hp(ids), _hp(out))
        return out

    def kernel_times(self) -> dict[str, float]:
        names = C.create_string_buffer(4096)
        us = (C.c_double * 64)()
        cnt = C.c_int()
        call("sph_sim_kernel_times", self._h, names, 4096, us, C.byref(cnt), 64)
        keys = names.value.decode().split(";") if cnt.value else []
        return {k: us[i] for i, k in enumerate(keys)}

    def pair_stats(self) -> dict[str, int]:
        v = [C.c_longlong() for _ in range(4)]
        call("sph_sim_pair_stats", self._h, *[C.byref(x) for x in v])
        return dict(zip(("candA", "hitsA", "candB", "hitsB"), (x.value for x in v)))

    @property
    def steps_done(self) -> int:
        s = C.c_longlong()
        call("sph_sim_steps_done", self._h, C.byref(s))
        return s.value
