"""Learned surrogate: a message-passing particle network (GNS-style) trained on
trajectories of the plain solver and run with the library's BF16 MFMA kernels.

* :class:`Gnn` -- ctypes wrapper of ``sph_gnn_*`` (forward, backward, Adam on
  the device).
* :func:`gnn_layout` -- the flat parameter layout (also returned by the library).
* :func:`record` / :func:`record_windows` / :func:`train` -- snapshots of
  plain 2-D dam-break runs, inputs = velocity history + wall distances + type
  + density, targets = mean acceleration and density rate over the next k
  steps (rates, so that one network serves cases of different dt).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import SphParams, call, load
from .sph import Case, Simulation

W = 128


class GnnCfg(C.Structure):
    _fields_ = [("layers", C.c_int), ("history", C.c_int), ("max_nodes", C.c_int), ("max_edges", C.c_int),
                ("train", C.c_int), ("seed", C.c_uint), ("v_scale", C.c_float), ("dv_scale", C.c_float),
                ("drho_scale", C.c_float), ("rho_scale", C.c_float), ("out_mean", C.c_float * 3),
                ("out_clip", C.c_float)]


_PROGRESS = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_void_p)


class RolloutStats(C.Structure):
    _fields_ = [("accepted", C.c_int), ("fallbacks", C.c_int), ("plain_steps", C.c_longlong),
                ("max_dev", C.c_double), ("mean_dev", C.c_double), ("seconds", C.c_double),
                ("rejected_penetration", C.c_int), ("max_penetration", C.c_double),
                ("rejected_density", C.c_int), ("rejected_graph", C.c_int), ("rejected_velocity", C.c_int),
                ("max_band_dev", C.c_double), ("max_bias_dev", C.c_double), ("plain_max_dev", C.c_double),
                ("plain_max_band_dev", C.c_double), ("plain_max_bias_dev", C.c_double), ("max_drift", C.c_double),
                ("skipped", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def gnn_layout(layers: int) -> np.ndarray:
    """Flat parameter layout: 5 ints per linear layer (w offset, N, K, b offset, 0)."""
    kins = [W, W] + [k for _ in range(layers) for k in (3 * W, 2 * W)] + [W]
    out, off = [], 0
    for kin in kins:
        for K in (kin, W, W):
            out.append((off, W, K, off + W * K, 0))
            off += W * K + W
    return np.asarray(out, np.int32).reshape(-1)


def _hp(a):
    return C.c_void_p(0) if a is None else a.ctypes.data_as(C.c_void_p)


class Gnn:
    """The surrogate network on the current device."""

    def __init__(self, layers=10, history=5, max_nodes=60_000, max_edges=1_500_000, train=True, seed=1,
                 v_scale=1.0, dv_scale=0.1, drho_scale=1.0, rho_scale=100.0, out_mean=(0.0, 0.0, 0.0),
                 out_clip=10.0):
        load()
        self.cfg = GnnCfg(layers, history, max_nodes, max_edges, 1 if train else 0, seed, v_scale, dv_scale,
                          drho_scale, rho_scale, (C.c_float * 3)(*out_mean), out_clip)
        h = C.c_void_p()
        call("sph_gnn_create", C.byref(self.cfg), C.byref(h))
        self._h = h
        n = C.c_longlong()
        call("sph_gnn_num_params", self._h, C.byref(n))
        self.nparams = n.value
        nl = C.c_int()
        lay = np.zeros(5 * (3 * (2 * layers + 3)), np.int32)
        call("sph_gnn_param_layout", self._h, _hp(lay), lay.size // 5, C.byref(nl))
        self.layout = lay[: 5 * nl.value]

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib._lib.sph_gnn_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        call("sph_gnn_get_params", self._h, _hp(out))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        call("sph_gnn_set_params", self._h, _hp(p))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        call("sph_gnn_get_grads", self._h, _hp(out))
        return out

    def load(self, params: SphParams, pos, hist, rho, typ, target=None) -> None:
        n = pos.shape[0]
        call("sph_gnn_load_state", self._h, C.byref(params), n, _hp(np.ascontiguousarray(pos, np.float32)),
             _hp(np.ascontiguousarray(hist, np.float32)), _hp(np.ascontiguousarray(rho, np.float32)),
             _hp(np.ascontiguousarray(typ, np.int32)),
             _hp(None if target is None else np.ascontiguousarray(target, np.float32)))
        self.n = n

    def graph(self):
        ne = C.c_int()
        call("sph_gnn_graph", self._h, C.byref(ne), None, None, None)
        off = np.empty(self.n + 1, np.int32)
        s = np.empty(ne.value, np.int32)
        r = np.empty(ne.value, np.int32)
        call("sph_gnn_graph", self._h, C.byref(ne), _hp(off), _hp(s), _hp(r))
        return off, s, r

    def features(self):
        off, s, _ = self.graph()
        xn = np.empty((self.n, W), np.float32)
        xe = np.empty((s.size, W), np.float32)
        call("sph_gnn_features", self._h, _hp(xn), _hp(xe))
        return xn, xe

    def set_fused(self, depth: int) -> None:
        """One fused kernel per MLP call with `depth` k-steps of loads in
        flight (1, 2 or 4; default for training contexts), three GEMM
        launches (0), 8 (inference contexts): persistent edge kernel with the
        first layer split into node projections, or 9 (the inference
        default): as 8 with BF16 edge latents."""
        call("sph_gnn_set_fused", self._h, int(depth))

    def tensor(self, kind: int, index: int = 0) -> np.ndarray:
        """An activation of the last forward pass (see sph_gnn_get_tensor):
        kinds 0/1 return BF16 bit patterns (uint16), the others FP32."""
        _, s, _ = self.graph()
        L = self.cfg.layers
        if kind in (0, 1):
            edges = index == 1 or (2 <= index < 2 + 2 * L and index % 2 == 0)
            a = np.empty(((s.size if edges else self.n), W), np.uint16)
        elif kind == 3:
            a = np.empty((s.size, W), np.float32)
        else:
            a = np.empty((self.n, W), np.float32)
        call("sph_gnn_get_tensor", self._h, int(kind), int(index), _hp(a), C.c_longlong(a.size))
        return a

    def forward(self, want_out=True):
        out = np.empty((self.n, 3), np.float32) if want_out else None
        loss = C.c_double()
        call("sph_gnn_forward", self._h, _hp(out), C.byref(loss))
        return out, loss.value

    def backward(self) -> None:
        call("sph_gnn_backward", self._h)

    def zero_grad(self) -> None:
        call("sph_gnn_zero_grad", self._h)

    def adam(self, lr: float, clip: float = 1.0) -> None:
        call("sph_gnn_adam", self._h, C.c_float(lr), C.c_float(clip))

    def rollout(self, sim: Simulation, k: int, intervals: int, dev_tol: float = 0.01, acc_tol: float = 0.5,
                drift_tol: float = 20.0, front_band: float = 0.02, band: float = 0.02, max_skip: int = 0,
                judge: Simulation | None = None, progress=None):
        """Advances `sim` by `intervals` x k steps with the surrogate, judged
        by the plain method's rates at both ends of every interval (evaluated
        in `judge`, a second context of the same case, made from the current
        state when not given) and backed up by the plain step; after a
        rejection the next 1, 2, 4, ... (at most max_skip; 0: none) intervals
        are plain without a network attempt.  Returns (stats, front, height).
        progress(done, accepted) is called every 1000 intervals."""
        f = np.zeros(max(intervals, 1))
        h = np.zeros(max(intervals, 1))
        st = RolloutStats()
        own = judge is None
        if own:
            cur = sim.state()
            judge = Simulation(Case(sim.params.copy(), cur["pos"], cur["vel"], cur["rho"], cur["p"], cur["type"],
                                    int((cur["type"] == 1).sum())))
        try:
            cb = _PROGRESS(lambda done, acc, _u: progress(done, acc)) if progress is not None else None
            call("sph_gnn_rollout", self._h, sim._h, judge._h, int(k), int(intervals), C.c_float(dev_tol),
                 C.c_float(acc_tol), C.c_float(drift_tol), C.c_double(front_band), C.c_double(band), int(max_skip),
                 _hp(f), _hp(h), C.byref(st),
                 C.cast(cb, C.c_void_p) if cb is not None else None, None)
        finally:
            if own:
                judge.close()
        return st.as_dict(), f[:intervals], h[:intervals]


def mfma_probe(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """D = A (16x32) @ B (32x16) through one gfx950 MFMA (inputs rounded to BF16)."""
    a = _to_bf16_bits(A)
    b = _to_bf16_bits(B)
    D = np.empty((16, 16), np.float32)
    call("sph_gnn_mfma_probe", _hp(a), _hp(b), _hp(D))
    return D


def _to_bf16_bits(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_round(x: np.ndarray) -> np.ndarray:
    return (_to_bf16_bits(x).astype(np.uint32) << 16).view(np.float32).reshape(np.shape(x))


# ------------------------------------------------------------------- data
@dataclass
class Snapshot:
    params: SphParams
    pos: np.ndarray     # (n, 3)
    hist: np.ndarray    # (history, n, 3), newest first
    rho: np.ndarray
    type: np.ndarray
    target: np.ndarray  # (n, 3) raw mean rates over the next k steps: dvx/dt, dvz/dt, drho/dt


@dataclass
class Dataset:
    k: int
    history: int
    snaps: list = field(default_factory=list)


def _rates(cur: dict, nxt: dict, T: float) -> np.ndarray:
    """Mean rates over an interval T between two frames: (dvx, dvz, drho) / T."""
    return (np.stack([nxt["vel"][:, 0] - cur["vel"][:, 0], nxt["vel"][:, 2] - cur["vel"][:, 2],
                      nxt["rho"] - cur["rho"]], 1).astype(np.float64) / T).astype(np.float32)


def record_batch(cases: list, nsteps: int, k: int, history: int) -> Dataset:
    """record() for several cases at once, stepped as one batched context."""
    ds = Dataset(k, history)
    frames = []
    with Simulation(cases) as sim:
        for i in range(nsteps // k + 1):
            frames.append(sim.state())
            if i < nsteps // k:
                sim.step(k)
        offs = sim.offsets
    for c, case in enumerate(cases):
        lo, hi = offs[c], offs[c + 1]
        fr = [{key: v[lo:hi] for key, v in f.items()} for f in frames]
        for t in range(history - 1, len(fr) - 1):
            cur, nxt = fr[t], fr[t + 1]
            hist = np.stack([fr[t - j]["vel"] for j in range(history)])
            tgt = _rates(cur, nxt, k * float(case.params.timeStep))
            ds.snaps.append(Snapshot(case.params, cur["pos"], hist.astype(np.float32), cur["rho"], cur["type"], tgt))
    return ds


def active(s: Snapshot) -> np.ndarray:
    """Fluid particles inside the escape box: the nodes of the loss (parked
    particles are out of the graph and the loss)."""
    P = s.params
    x = s.pos[:, 0].astype(np.float64)
    z = s.pos[:, 2].astype(np.float64)
    out = (x > P.domain_max[0]) | (x < P.domain_min[0]) | (z > P.domain_max[2]) | (z < P.domain_min[2])
    return (s.type == 1) & ~out


def normalisation(snaps: list) -> dict:
    """Output scales and means over the active fluid particles of the snapshots."""
    tg = np.concatenate([s.target[active(s)] for s in snaps]).astype(np.float64)
    mean = tg.mean(0)
    std = tg.std(0) + 1e-9
    return {"dv_scale": float(max(std[0], std[1])), "drho_scale": float(std[2]),
            "out_mean": (float(mean[0]), float(mean[1]), float(mean[2]))}


def normalised_target(s: Snapshot, norm: dict) -> np.ndarray:
    m = np.asarray(norm["out_mean"], np.float64)
    sc = np.array([norm["dv_scale"], norm["dv_scale"], norm["drho_scale"]])
    return ((s.target - m) / sc).astype(np.float32)


def train(g: "Gnn", snaps: list, norm: dict, iters: int, batch: int = 4, lr: float = 2e-4,
          lr_final: float = 2e-5, clip: float = 1.0, seed: int = 0, seconds: float | None = None, progress=None,
          target_clip: float | None = 10.0) -> list:
    """Adam on minibatches of `batch` snapshots (gradients summed over the
    batch by the backward pass, clipped at global norm clip x batch, i.e. clip
    for the mean gradient); cosine learning-rate decay over the iterations.
    Snapshot order comes from `seed` alone, so a run is reproduced by its seed
    and iteration count (`seconds`, if given, only stops a run early).
    progress(iterations, recent mean loss) is called every 30 s.  Normalised
    training targets are clipped to +-target_clip: the few splash particles
    whose rates lie tens of standard deviations out would otherwise dominate
    the batches (evaluation uses the unclipped targets)."""
    import time
    rng = np.random.default_rng(seed)
    targets = []
    for s in snaps:
        t = normalised_target(s, norm)
        targets.append(np.clip(t, -target_clip, target_clip) if target_clip is not None else t)
    batch = max(1, min(batch, len(snaps)))
    losses = []
    t0 = t_last = time.time()
    for it in range(iters):
        g.zero_grad()
        loss = 0.0
        for j in rng.choice(len(snaps), size=batch, replace=False):
            s = snaps[j]
            g.load(s.params, s.pos, s.hist, s.rho, s.type, targets[j])
            _, lj = g.forward(want_out=False)
            g.backward()
            loss += lj / batch
        frac = it / max(iters - 1, 1)
        g.adam(lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac)), clip * batch)
        losses.append(loss)
        if progress is not None and time.time() - t_last > 30.0:
            t_last = time.time()
            progress(it + 1, float(np.mean(losses[-50:])))
        if seconds is not None and time.time() - t0 > seconds:
            break
    return losses


def save_weights(path, g: "Gnn", norm: dict, k: int, meta: dict | None = None) -> None:
    """Trained network: FP32 master parameters, the configuration that
    defines its inputs and outputs, the interval k it was trained for and
    free-form metadata (JSON)."""
    import json
    c = g.cfg
    np.savez_compressed(path, params=g.params().astype(np.float32), layers=c.layers, history=c.history, k=k,
                        v_scale=c.v_scale, rho_scale=c.rho_scale, dv_scale=norm["dv_scale"],
                        drho_scale=norm["drho_scale"], out_mean=np.asarray(norm["out_mean"], np.float64),
                        out_clip=c.out_clip, meta=json.dumps(meta or {}))


def load_weights(path) -> dict:
    """Reads save_weights' file: {"params", "k", "meta", "cfg": Gnn keyword
    arguments (layers, history, scales, out_mean, out_clip)}."""
    import json
    with np.load(path, allow_pickle=False) as z:
        cfg = {"layers": int(z["layers"]), "history": int(z["history"]), "v_scale": float(z["v_scale"]),
               "rho_scale": float(z["rho_scale"]), "dv_scale": float(z["dv_scale"]),
               "drho_scale": float(z["drho_scale"]), "out_mean": tuple(float(v) for v in z["out_mean"]),
               "out_clip": float(z["out_clip"])}
        return {"params": z["params"].astype(np.float32), "k": int(z["k"]), "cfg": cfg,
                "meta": json.loads(str(z["meta"]))}


def judge_interval(judge: Simulation, k: int, s0: dict, s1: dict, front_band: float = 0.02) -> dict:
    """The rollout's judge on one interval of k steps from state s0 to s1
    (dicts of pos, vel, rho, p, type as Simulation.state() returns them);
    `judge` is a context of the case (its state is replaced)."""
    out = np.zeros(7)
    arr = lambda a, d: np.ascontiguousarray(a, d)  # noqa: E731
    call("sph_gnn_judge", judge._h, int(k), C.c_double(front_band), _hp(arr(s0["pos"], np.float32)),
         _hp(arr(s0["vel"], np.float32)), _hp(arr(s0["rho"], np.float32)), _hp(arr(s0["p"], np.float32)),
         _hp(arr(s0["type"], np.int32)), _hp(arr(s1["pos"], np.float32)), _hp(arr(s1["vel"], np.float32)),
         _hp(arr(s1["rho"], np.float32)), _hp(arr(s1["p"], np.float32)), _hp(out))
    return {"max_dev": out[0], "mean_dev": out[1], "band_dev": out[2], "bias_dev": out[3], "escaped": int(out[4]),
            "max_penetration": out[5], "band_count": int(out[6])}


def evaluate(g: "Gnn", snaps: list, norm: dict) -> dict:
    """Held-out mean squared error of the normalised outputs, next to the
    errors of the method without the network: predicting the mean change
    (constant acceleration) and predicting no change (constant velocity)."""
    net, mean_pred, zero_pred = [], [], []
    m = np.asarray(norm["out_mean"], np.float64)
    sc = np.array([norm["dv_scale"], norm["dv_scale"], norm["drho_scale"]])
    for s in snaps:
        t = normalised_target(s, norm)
        g.load(s.params, s.pos, s.hist, s.rho, s.type, t)
        _, loss = g.forward(want_out=False)
        fl = active(s)
        net.append(loss)
        mean_pred.append(float(np.mean(t[fl].astype(np.float64) ** 2)))
        zero_pred.append(float(np.mean(((0.0 - m) / sc - t[fl]) ** 2)))
    return {"network": float(np.mean(net)), "constant_acceleration": float(np.mean(mean_pred)),
            "constant_velocity": float(np.mean(zero_pred))}


def record(case: Case, nsteps: int, k: int, history: int) -> Dataset:
    """Runs the plain solver and keeps a snapshot every k steps (with history)."""
    ds = Dataset(k, history)
    frames = []
    with Simulation(case) as sim:
        for i in range(nsteps // k + 1):
            st = sim.state()
            frames.append(st)
            if i < nsteps // k:
                sim.step(k)
    for t in range(history - 1, len(frames) - 1):
        cur, nxt = frames[t], frames[t + 1]
        hist = np.stack([frames[t - j]["vel"] for j in range(history)])
        tgt = _rates(cur, nxt, k * float(case.params.timeStep))
        ds.snaps.append(Snapshot(case.params, cur["pos"], hist.astype(np.float32), cur["rho"], cur["type"], tgt))
    return ds


def record_windows(cases: list, total_steps: int, k: int, history: int, samples: int, progress=None) -> Dataset:
    """Snapshots of one batched plain run of `cases` over `total_steps` steps:
    at `samples` evenly spaced sample steps T the frames T - (history-1) k,
    ..., T give the velocity history and the frame T + k the target.  Only the
    frames of the windows are downloaded, so long runs stay cheap."""
    ds = Dataset(k, history)
    span = history * k  # steps covered by one window
    first = (history - 1) * k
    if samples < 1 or total_steps < first + k:
        raise ValueError("run too short for one window")
    gap = max((total_steps - first - k) // max(samples - 1, 1), span)
    with Simulation(cases) as sim:
        sim.capture()
        cur = 0
        offs = sim.offsets
        for s in range(samples):
            T = first + s * gap
            if T + k > total_steps:
                break
            start = T - first
            if start > cur:
                sim.replay(start - cur)
                cur = start
            frames = []
            for h in range(history + 1):
                frames.append(sim.state())
                if h < history:
                    sim.replay(k)
                    cur += k
            for c, case in enumerate(cases):
                lo, hi = offs[c], offs[c + 1]
                fr = [{key: v[lo:hi] for key, v in f.items()} for f in frames]
                now, nxt = fr[history - 1], fr[history]
                hist = np.stack([fr[history - 1 - j]["vel"] for j in range(history)])
                tgt = _rates(now, nxt, k * float(case.params.timeStep))
                ds.snaps.append(Snapshot(case.params, now["pos"], hist.astype(np.float32), now["rho"], now["type"],
                                         tgt))
            if progress is not None:
                progress(s + 1, cur)
    return ds


def forward_flops(layers: int, n: int, edges: int) -> float:
    """Multiply-add work (2 flop each) of one forward pass: three 128-wide
    linear layers per MLP; encoder and processor inputs are 128 (encoders),
    384 (edge update) and 256 (node update) wide."""
    per_row = lambda kin: 2.0 * W * (kin + W + W)
    return (n * per_row(W) + edges * per_row(W) + layers * (edges * per_row(3 * W) + n * per_row(2 * W))
            + n * per_row(W))


def learned_vs_plain(g: "Gnn", case: Case, k: int, t_end: float = 2.0, dev_tol: float = 0.01,
                     acc_tol: float = 0.5, drift_tol: float = 20.0, front_band: float = 0.02, band: float = 0.02,
                     max_skip: int = 0, tuning: dict | None = None, progress=None) -> dict:
    """Plain GPU solver against the surrogate rollout on one case, both to
    t_end.  The plain run replays its captured step and probes the surge front
    every k steps on the device (one host wait at the end), timed as a whole.
    The surrogate rollout (k plain steps per network step, judged by the plain
    rates at both ends of every interval, plain steps as the fallback) runs
    over the same intervals, timed as a whole, and probes the front after
    every interval; the front error is taken at every interval."""
    import time
    dt = float(case.params.timeStep)
    intervals = max(int(round(t_end / (dt * k))), 1)
    with Simulation(case) as sim:
        if tuning:
            sim.set_tuning(**tuning)
        sim.capture()
        sim.synchronize()
        t0 = time.perf_counter()
        front_p, _ = sim.probe_history(k, intervals, band)
        plain_s = time.perf_counter() - t0
        tuning_used = sim.tuning.as_dict()
    with Simulation(case) as sim, Simulation(case) as judge:
        if tuning:
            sim.set_tuning(**tuning)
        st, front_l, _ = g.rollout(sim, k, intervals, dev_tol=dev_tol, acc_tol=acc_tol, drift_tol=drift_tol,
                                   front_band=front_band, band=band, max_skip=max_skip, judge=judge,
                                   progress=progress)
    rel = np.abs(front_l - front_p) / np.maximum(front_p, 1e-9)
    worst = int(np.argmax(rel))
    return {"particles": int(case.n), "fluid": int(case.n_fluid), "dt": dt, "k": k, "intervals": intervals,
            "t_end": intervals * k * dt, "plain_seconds": plain_s, "learned_seconds": st["seconds"],
            "speedup": plain_s / st["seconds"], "tuning": tuning_used,
            "dev_tol": dev_tol, "acc_tol": acc_tol, "drift_tol": drift_tol, "front_band": front_band,
            "max_skip": max_skip,
            "front_rel_error_max": float(rel.max()), "front_rel_error_max_at_t": float((worst + 1) * k * dt),
            "front_rel_error_final": float(rel[-1]),
            "front_plain_final": float(front_p[-1]), "front_learned_final": float(front_l[-1]),
            "accepted_fraction": st["accepted"] / max(intervals, 1), "rollout": st}
