"""Benchmark of the delta-SPH dam-break step on one MI355X.

Workload: the 3-D dam break of 2.0M fluid particles -- a 1.0 x 0.5 x 0.5 m
column in a 3.2 x 1.0 x 1.0 m tank, dx = 0.005 m, 3 layers of dummy wall
particles (1.41M), Wendland C2 with h = 1.3 dx, Tait EOS, delta = 0.1,
alpha = 0.02, CFL 0.2.  A step is one full predictor-corrector step
(2 x hash, sort, cell ranges, gather, kernel A, kernel B, integrator)
replayed from a captured hipGraph with the state resident in HBM.

Prints one JSON line: particle-steps/s (fluid particles x steps / s), the
roofline of the dominant kernel measured with HIP events on its stream, and a
CPU baseline (the oracle's FP32 step on the host cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0


def _workload(name: str):
    """(description, list of cases)."""
    from tmp_sph_amd import cases, sph
    if name == "c4":
        cs = cases.sweep_cases()
        return (f"batched sweep: {len(cs)} uneven 2-D dam breaks (H~U[0.2,0.6] m, W/H~U[0.5,2], "
                f"dx in {{0.002,0.004}} m, 4.0x1.0 m tank) in one graph per step", cs)
    desc, cfg = cases.named(name)
    return desc, [sph.make_case(cfg)]


def _algorithmic_bytes(name: str, dim: int, n: int, nf: int) -> float:
    """Compulsory HBM bytes of one launch of a kernel (each array touched once).

    nsB: PR (16 B) of every target; VP (16 B), dofv (4 B) and key (4 B) of
    fluid targets; RATE (16 B) written for every target.  nsA: PR, VP, key read
    (36 B) and dofv written (4 B) for every target, wall p and rho (8 B).
    nbrlist (neighbour-list mode): PR and key read, count written.
    """
    nw = n - nf
    if name == "nbrlist":  # PR and key read, count written (the lists themselves are intermediate)
        return 24.0 * n
    if name == "nsB":
        return 16.0 * n + 24.0 * nf + 16.0 * n
    if name == "nsA":
        return 40.0 * n + 8.0 * nw
    raise KeyError(name)


def cpu_model() -> str:
    """The host CPU's model name (lscpu), for the baseline's record."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    import platform
    return platform.processor() or "unknown"


ORACLE_BUILD = "g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp, FP32 fields (oracle mode 0)"


def _advanced_case(case, state):
    out = case.copy()
    out.pos, out.vel, out.rho, out.p, out.type = state["pos"], state["vel"], state["rho"], state["p"], state["type"]
    return out


def cpu_rate(case, steps: int, runs: int, threads: int, warmup: int = 0) -> dict:
    """The oracle's FP32 step on `threads` host threads: `runs` runs of
    `steps` steps each (the step loop timed, steady_clock-style wall time),
    started after `warmup` untimed steps from rest; median fluid
    particle-steps/s."""
    from oracle import oracle as orc
    import numpy as np
    start = case
    if warmup > 0:  # the warm-up steps on every thread: they only set the start state
        st, _, _, _ = orc.run(case.params, case, warmup, mode=0)
        start = _advanced_case(case, st)
    old = orc.num_threads(0)
    orc.num_threads(threads)
    secs = []
    try:
        for r in range(runs):
            secs.append(orc.run(start.params, start, steps, mode=0)[3])
            print(f"#   {case.n} particles, {threads} thread(s): run {r + 1} of {runs}, {steps} steps in "
                  f"{secs[-1]:.2f} s", file=sys.stderr, flush=True)
    finally:
        orc.num_threads(old)
    rates = [case.n_fluid * steps / s for s in secs]
    return {"value": float(np.median(rates)), "runs": rates, "steps_per_run": steps, "warmup_steps": warmup,
            "threads": threads, "seconds": secs}


def cpu_baseline(case, seconds_target: float) -> dict:
    """Bounded sample of the same workload on the host cores: all threads
    (median of 5 runs) and one thread (one step), with the CPU model."""
    from oracle import oracle as orc
    threads = orc.num_threads(0)
    _, _, _, s1 = orc.run(case.params, case, 1, mode=0)  # also the first-touch warm-up
    k = max(1, min(20, int(seconds_target / 5 / max(s1, 1e-9))))
    allc = cpu_rate(case, k, 5, threads)
    one = cpu_rate(case, 1, 1, 1)
    return {"value": allc["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 reference step (C++ + OpenMP) on {threads} threads, median of 5 runs of {k} "
                      f"step(s) from rest, {case.n} particles ({case.n_fluid} fluid)",
            "runs": allc["runs"], "serial": {"value": one["value"], "cores": 1, "steps": 1,
                                             "seconds": one["seconds"][0]},
            "cpu_model": cpu_model(), "build": ORACLE_BUILD}


def cpu_plan(args) -> int:
    """The host baseline table of every plain configuration: C1 (2000 steps),
    C2 and C3 (20 steps after 5 warm-up steps), C4 (every sweep member, 20
    steps after 5), all threads and one thread, median of 5 runs where the
    time allows; one JSON line."""
    from oracle import oracle as orc
    from tmp_sph_amd import cases as cs, sph
    threads = orc.num_threads(0)
    out = {"metric": "host baseline (oracle FP32 step)", "unit": "fluid particle-steps/s", "cpu_model": cpu_model(),
           "threads_all": threads, "build": ORACLE_BUILD, "configs": {}}

    def row(name, case, steps, warmup, runs_all, runs_one, steps_one=None):
        t0 = time.time()
        r = {"particles": case.n, "fluid": case.n_fluid,
             "all": cpu_rate(case, steps, runs_all, threads, warmup),
             "one": cpu_rate(case, steps_one or steps, runs_one, 1, warmup)}
        out["configs"][name] = r
        print(f"# {name}: {r['all']['value']:.4g} ({threads} threads), {r['one']['value']:.4g} (1 thread) "
              f"fluid particle-steps/s ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)

    c1 = sph.make_case(cs.named("c1")[1])
    row("c1", c1, 2000, 0, 5, 5)
    c2 = sph.make_case(cs.named("c2")[1])
    row("c2", c2, 20, 5, 5, 1)
    c3 = sph.make_case(cs.named("c3")[1])
    row("c3", c3, 20, 5, 5, 1, steps_one=3)
    t0 = time.time()
    secs, fl = 0.0, 0
    for m, c in enumerate(cs.sweep_cases()):
        st, _, _, _ = orc.run(c.params, c, 5, mode=0)
        secs += orc.run(c.params, _advanced_case(c, st), 20, mode=0)[3]
        fl += c.n_fluid
        if (m + 1) % 8 == 0:
            print(f"#   c4: {m + 1} members done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    out["configs"]["c4"] = {"instances": len(cs.sweep_cases()), "fluid": fl,
                            "all": {"value": fl * 20 / secs, "steps_per_run": 20, "warmup_steps": 5,
                                    "threads": threads, "seconds": secs,
                                    "note": "members run one after another (the oracle has no batched mode); "
                                            "one run"}}
    print(f"# c4: {fl * 20 / secs:.4g} fluid particle-steps/s ({time.time() - t0:.0f} s)", file=sys.stderr,
          flush=True)
    print(json.dumps(out), flush=True)
    return 0


def copy_probe(mib: int) -> int:
    """A plain device copy of known size, for converting the TCC byte
    counters (FETCH_SIZE, WRITE_SIZE) of a kernel into HBM bytes."""
    from tmp_sph_amd import _lib, sph
    n = mib << 20
    a = sph.DeviceArray(n, "u1")
    b = sph.DeviceArray(n, "u1")
    a.fill_bytes(1)
    for _ in range(5):
        _lib.call("sph_memcpy_dtod", b.c, a.c, n)
    _lib.call("sph_device_synchronize")
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 5}), flush=True)
    return 0


def output_bench(sim, every: int, frames: int, ms_plain: float) -> dict:
    """Wall time of `frames` x (`every` replayed steps + one binary VTK frame):
    with the asynchronous writer (the frame's copy and file write overlap the
    next steps) and with a blocking download followed by the same write."""
    import shutil
    import tempfile
    from tmp_sph_amd import files
    tmp = Path(tempfile.mkdtemp(prefix="sph_frames_"))
    try:
        sim.output_open(tmp / "async", binary=True)
        sim.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            sim.replay(every)
            sim.output_frame(f)
        sim.output_flush()
        sim.synchronize()
        t_async = time.perf_counter() - t0
        sim.output_close()
        (tmp / "sync").mkdir()
        t0 = time.perf_counter()
        for f in range(frames):
            sim.replay(every)
            st = sim.state()
            files.write_vtk(tmp / "sync" / f"fluid_t={f}.vtk", st["pos"], st["vel"], st["p"], binary=True)
        sim.synchronize()
        t_sync = time.perf_counter() - t0
        frame_bytes = (tmp / "sync" / "fluid_t=0.vtk").stat().st_size
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    steps = every * frames
    res = {"every": every, "frames": frames, "frame_bytes": frame_bytes,
           "ms_per_step_no_output": ms_plain,
           "ms_per_step_async": 1000.0 * t_async / steps, "ms_per_step_blocking": 1000.0 * t_sync / steps}
    print(f"# output every {every} steps: {res['ms_per_step_async']:.3f} ms/step asynchronous, "
          f"{res['ms_per_step_blocking']:.3f} blocking, {ms_plain:.3f} without frames", file=sys.stderr, flush=True)
    return res


def surrogate_bench(args) -> int:
    """Throughput of the surrogate network's BF16 MFMA kernels on the graph of
    the learned benchmark's evaluation case (0.8 x 0.5 m column, dx = 0.002
    m: 109 k particles, 2.1 M edges): inference forward with the fused MLP
    kernel and with three GEMM launches per MLP, and one training step
    (forward, backward, Adam) on the dx = 0.004 m version (29.5 k particles)."""
    import numpy as np
    from tmp_sph_amd import sph
    from tmp_sph_amd import surrogate as sg
    layers = 10
    reps = max(1, args.steps)

    def timed(f, n):
        f()
        sph.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            f()
        sph.synchronize()
        return (time.perf_counter() - t0) / n

    case = sph.make_case(sph.dambreak_cfg(dim=2, dx=0.002, tank=(4.0, 1.0, 1.0), column=(0.8, 1.0, 0.5)))
    hist = (np.random.default_rng(0).standard_normal((5, case.n, 3)) * 0.1).astype(np.float32)
    g = sg.Gnn(layers=layers, max_nodes=case.n, max_edges=30 * case.n, train=False)
    g.load(case.params, case.pos, hist, case.rho, case.type)
    _, send, _ = g.graph()
    ne = int(send.size)
    flops = sg.forward_flops(layers, case.n, ne)
    res = {}
    for depth in (9, 8, 2, 0):
        g.set_fused(depth)
        dt = timed(lambda: g.forward(want_out=False), reps)
        name = {9: "persistent_split_bf16_latents", 8: "persistent_split", 0: "three_gemms"}.get(
            depth, f"fused_depth{depth}")
        res[name] = {"ms": 1e3 * dt, "tflops": flops / dt / 1e12}
        print(f"# forward ({name}): {1e3 * dt:.2f} ms, {flops / dt / 1e12:.1f} TFLOP/s", file=sys.stderr, flush=True)
    g.close()
    small = sph.make_case(sph.dambreak_cfg(dim=2, dx=0.004, tank=(4.0, 1.0, 1.0), column=(0.8, 1.0, 0.5)))
    h2 = (np.random.default_rng(1).standard_normal((5, small.n, 3)) * 0.1).astype(np.float32)
    tg = np.random.default_rng(2).standard_normal((small.n, 3)).astype(np.float32)
    gt = sg.Gnn(layers=layers, max_nodes=small.n, max_edges=30 * small.n, train=True)
    gt.load(small.params, small.pos, h2, small.rho, small.type, tg)
    _, s2, _ = gt.graph()

    def train_step():
        gt.zero_grad()
        gt.forward(want_out=False)
        gt.backward()
        gt.adam(1e-5)

    dtt = timed(train_step, max(1, reps // 2))
    tflops_train = 3 * sg.forward_flops(layers, small.n, int(s2.size)) / dtt / 1e12
    print(f"# training step: {1e3 * dtt:.2f} ms (~{tflops_train:.1f} TFLOP/s counting 3x forward flops)",
          file=sys.stderr, flush=True)
    gt.close()
    best = res["persistent_split_bf16_latents"]
    out = {
        "metric": "surrogate forward throughput (edges x layers / s)",
        "value": ne * layers / (best["ms"] * 1e-3),
        "unit": "edge-layers/s",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": best["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 MFMA, fp32 accumulate",
        "data": "synthetic: random velocity history, random-init weights",
        "config": {"workload": "message-passing network, 10 blocks, latent 128, on the 2.1 M-edge graph of a "
                               "0.8x0.5 m column at dx=0.002 m", "nodes": case.n, "edges": ne},
        "roofline": {"bound": "mfma", "achieved": best["tflops"], "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": best["tflops"] / BF16_MFMA_PEAK_TFLOPS, "traffic": None},
        "forward": res,
        "training_step": {"nodes": small.n, "edges": int(s2.size), "ms": 1e3 * dtt, "tflops_3x_forward": tflops_train},
    }
    print(json.dumps(out), flush=True)
    return 0


# steps of each plain configuration's whole run (the default timed length)
RUN_STEPS = {"c1": 2000, "c2": 20000, "c3": 10000, "c4": 5000}

WEIGHTS = ROOT / "tmp_sph_amd" / "weights" / "surrogate_k10_l10.npz"
# held-out evaluation geometry: a 0.8 x 0.5 m column in the 4.0 x 1.0 m tank at
# dx = 0.002 m (100 k fluid particles; no sweep member has this column)
C5_CFG = dict(dim=2, dx=0.002, tank=(4.0, 1.0, 1.0), column=(0.8, 1.0, 0.5))
# judge thresholds: density deviation 1 % of rho0 per particle; front-band and
# whole-flow mean velocity residuals 0.5 g T per network step (plain intervals
# of this case read up to 0.15 g T in the front band while the column starts
# to collapse); front-band residual summed over the accepted steps 20 g T
# (0.033 m/s of front velocity at this case's T = 10 dt)
LEARNED_TOL = {"dev_tol": 0.01, "acc_tol": 0.5, "drift_tol": 20.0, "front_band": 0.02}
# after a rejected network step, the next 1, 2, 4, ... up to this many
# intervals are plain without a network attempt (0: attempt every interval)
LEARNED_MAX_SKIP = 64


def learned_eval(weights: Path, t_end: float = 2.0, tol: dict | None = None,
                 max_skip: int = LEARNED_MAX_SKIP) -> dict:
    """The trained surrogate against the plain GPU solver on the held-out
    geometry, both run to t_end (no extrapolation in time), front error at
    every interval, plus the speed ceiling: the plain time per interval
    against the network's forward time (measured, and at the BF16 MFMA peak)."""
    import numpy as np
    from tmp_sph_amd import sph
    from tmp_sph_amd import surrogate as sg
    tol = dict(LEARNED_TOL if tol is None else tol)
    w = sg.load_weights(weights)
    case = sph.make_case(sph.dambreak_cfg(**C5_CFG))
    g = sg.Gnn(max_nodes=case.n, max_edges=30 * case.n, train=False, **w["cfg"])
    g.set_params(w["params"])
    t0 = time.time()
    print(f"# learned vs plain on {case.n} particles to t = {t_end} s: plain run first", file=sys.stderr, flush=True)
    res = sg.learned_vs_plain(g, case, w["k"], t_end, **tol, max_skip=max_skip, progress=lambda done, acc: print(
        f"# learned rollout: {done} intervals, {acc} accepted ({time.time() - t0:.0f} s)", file=sys.stderr,
        flush=True))
    print(f"# learned vs plain to t = {res['t_end']:.3f} s: {res['speedup']:.3f}x, front error max "
          f"{res['front_rel_error_max']:.4f}, accepted {res['rollout']['accepted']} of {res['intervals']} "
          f"({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    hist = np.zeros((w["cfg"]["history"], case.n, 3), np.float32)
    g.load(case.params, case.pos, hist, case.rho, case.type)
    _, send, _ = g.graph()
    g.forward(want_out=False)
    reps = 5
    sph.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        g.forward(want_out=False)
    fwd_s = (time.perf_counter() - t1) / reps
    flops = sg.forward_flops(w["cfg"]["layers"], case.n, int(send.size))
    g.close()
    plain_iv = res["plain_seconds"] / res["intervals"]
    peak_s = flops / (BF16_MFMA_PEAK_TFLOPS * 1e12)
    return {
        "speedup": res["speedup"], "front_rel_error_max": res["front_rel_error_max"], "t_end": res["t_end"],
        "front_within_2pct": bool(res["front_rel_error_max"] <= 0.02),
        "unit": "x (plain wall time / learned wall time to t_end, same intervals)",
        "workload": f"held-out 2-D dam break, 0.8x0.5 m column in 4.0x1.0 m tank, dx=0.002 m "
                    f"({case.n} particles, {case.n_fluid} fluid)",
        "network": f"k={w['k']} plain steps per network step, {w['cfg']['layers']} message-passing blocks, "
                   f"latent 128, BF16 MFMA with FP32 accumulation",
        "weights": str(Path(weights).relative_to(ROOT)) if Path(weights).is_relative_to(ROOT) else str(weights),
        "judge": tol, "max_skip": max_skip,
        "accepted": res["rollout"]["accepted"], "intervals": res["intervals"],
        "plain_seconds": res["plain_seconds"], "learned_seconds": res["learned_seconds"],
        "ceiling": {"plain_ms_per_interval": 1e3 * plain_iv, "forward_ms": 1e3 * fwd_s,
                    "forward_tflops": flops / fwd_s / 1e12, "mfma_frac": flops / fwd_s / 1e12 / BF16_MFMA_PEAK_TFLOPS,
                    "speedup_if_all_accepted": plain_iv / fwd_s,
                    "forward_ms_at_mfma_peak": 1e3 * peak_s, "speedup_if_all_accepted_at_mfma_peak": plain_iv / peak_s},
        "evaluation": res,
    }


def reference_surface_bench(args) -> int:
    """The step through the drop-in reference call surface on the C3 case:
    per half-step calcHash, sortParticles, reorder (v0 permutation table),
    SPH_NS_simpleversion and PC_prediction / PC_correction on caller-owned
    device arrays in the reference's AoS layouts, in the order of
    opt_src_v0/sph/sph_run.cu:269-309 -- what a caller gets by rebinding the
    reference's own calls, next to the default line's fast-path context."""
    import numpy as np
    from tmp_sph_amd import cases, sph
    desc, cfg = cases.named("c3")
    case = sph.make_case(cfg)
    prm = case.params
    n = case.n
    sph.set_params(prm)
    sph.reserve_scratch(n)
    D, E = sph.DeviceArray.from_host, sph.DeviceArray
    S = [{"pos": D(case.pos), "vel": D(case.vel), "rho": D(case.rho), "p": D(case.p), "type": D(case.type)},
         {"pos": E((n, 3), np.float32), "vel": E((n, 3), np.float32), "rho": E((n,), np.float32),
          "p": E((n,), np.float32), "type": E((n,), np.int32)}]
    tabs = [D(np.zeros(n, np.int32)), D(np.zeros(n, np.int32))]
    h, idx = E((n,), np.int32), E((n,), np.int32)
    cs, ce = E((prm.numCells,), np.int32), E((prm.numCells,), np.int32)
    drho, acc, dofv = E((n,), np.float32), E((n, 3), np.float32), E((n,), np.float32)
    tpos, tvel, trho = E((n, 3), np.float32), E((n, 3), np.float32), E((n,), np.float32)

    def half(a, b, pred):
        A, B = S[a], S[b]
        sph.calcHash(A["pos"], h, idx, n)
        sph.sortParticles(h, idx, n)
        sph.reorder(A["pos"], A["vel"], A["rho"], A["p"], A["type"], None, B["pos"], B["vel"], B["rho"], B["p"],
                    B["type"], None, h, idx, cs, ce, n, tabs[a], tabs[b])
        sph.SPH_NS_simpleversion(B["pos"], B["vel"], B["rho"], B["p"], B["type"], drho, acc, cs, ce, n, h, 0, dofv)
        f = sph.PC_prediction if pred else sph.PC_correction
        f(B["pos"], B["vel"], B["rho"], B["p"], drho, acc, h, idx, tpos, tvel, trho, n, B["type"])

    def step():
        half(0, 1, True)
        half(1, 0, False)

    steps = args.steps if args.steps > 0 else 200
    for _ in range(args.warmup):
        step()
    sph.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sph.synchronize()
    el = time.perf_counter() - t0
    print(f"# reference call surface: {steps} steps, {1e3 * el / steps:.3f} ms/step", file=sys.stderr, flush=True)
    out = {"metric": "particle-steps/s through the reference call surface (drop-in)",
           "value": case.n_fluid * steps / el, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * el / steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic dam-break lattice released from rest",
           "config": {"workload": desc + ", through calcHash / sortParticles / reorder / SPH_NS_simpleversion / "
                                         "PC_prediction / PC_correction on AoS device arrays",
                      "particles_total": n, "particles_fluid": case.n_fluid}}
    print(json.dumps(out), flush=True)
    return 0


def learned_bench(args) -> int:
    """Trains the surrogate and evaluates it against the plain GPU solver.

    Training data: windows of one batched plain run (to t = 2 s) of the first
    --train-members sweep members with dx = 0.002 m (velocity history of 5
    frames k steps apart, target = mean rates over the next k steps; k = 10
    unless --k says otherwise).
    Training: Adam on minibatches of 4 snapshots on the GPU, --train-iters
    iterations, snapshot order from --seed.  Saves the weights
    (--save-weights) and evaluates them with learned_eval to --learned-t."""
    import numpy as np
    from tmp_sph_amd import cases as cs
    from tmp_sph_amd import surrogate as sg
    k, history, layers = args.k, 5, 10
    t0 = time.time()
    members = [c for c, m in zip(cs.sweep_cases(), cs.sweep_members()) if m.dx == 0.002][:args.train_members]
    dts = sorted(float(c.params.timeStep) for c in members)
    total = int(2.0 / dts[len(dts) // 2])
    ds = sg.record_windows(members, total, k, history, samples=30, progress=lambda smp, step: print(
        f"# recording: window {smp} at step {step} ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True))
    print(f"# recorded {len(ds.snaps)} snapshots of {len(members)} instances over {total} steps "
          f"({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(0)
    order = rng.permutation(len(ds.snaps))
    held = [ds.snaps[i] for i in order[:24]]
    train = [ds.snaps[i] for i in order[24:]]
    norm = sg.normalisation(train)
    nmax = max(s.pos.shape[0] for s in ds.snaps)
    kw = dict(layers=layers, history=history, v_scale=3.0, rho_scale=100.0, dv_scale=norm["dv_scale"],
              drho_scale=norm["drho_scale"], out_mean=norm["out_mean"])
    g = sg.Gnn(max_nodes=nmax, max_edges=30 * nmax, train=True, seed=11 + args.seed, **kw)
    before = sg.evaluate(g, held, norm)
    t1 = time.time()
    losses = sg.train(g, train, norm, iters=args.train_iters, batch=4, seed=args.seed, seconds=args.train_seconds,
                      progress=lambda it, l: print(f"# training: {it} iterations, loss {l:.3f}", file=sys.stderr,
                                                   flush=True))
    train_s = time.time() - t1
    after = sg.evaluate(g, held, norm)
    print(f"# trained {len(losses)} iterations in {train_s:.0f} s: held-out mse {before['network']:.3f} -> "
          f"{after['network']:.3f} (constant-acceleration baseline {after['constant_acceleration']:.3f})",
          file=sys.stderr, flush=True)
    save = Path(args.save_weights) if args.save_weights else ROOT / "gpurun_out" / f"surrogate_seed{args.seed}.npz"
    save.parent.mkdir(parents=True, exist_ok=True)
    meta = {"seed": args.seed, "iterations": len(losses), "batch": 4, "train_seconds": train_s,
            "members": [m.k for m in cs.sweep_members() if m.dx == 0.002][:args.train_members],
            "snapshots": len(train), "held_out_mse_before": before, "held_out_mse_after": after,
            "loss_first_50": float(np.mean(losses[:50])), "loss_last_50": float(np.mean(losses[-50:]))}
    sg.save_weights(save, g, norm, k, meta)
    g.close()
    print(f"# weights saved to {save}", file=sys.stderr, flush=True)
    learned = learned_eval(save, args.learned_t, max_skip=args.learned_max_skip) if args.learned_t > 0 else None
    out = {
        "metric": "learned vs plain speedup at <=2% front error",
        "value": learned["speedup"] if learned else None,
        "unit": "x (plain wall time / learned wall time to t = 2 s)",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "bf16 MFMA (fp32 accumulate) network, f32 solver",
        "data": f"synthetic: trained on plain-solver trajectories of {len(members)} sweep members (dx = 0.002 m)",
        "config": {"workload": "held-out 2-D dam break, 0.8x0.5 m column in 4.0x1.0 m tank, dx=0.002 m",
                   "k_steps_per_net_step": k, "layers": layers, "latent": 128},
        "training": meta,
        "learned": learned,
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default: the configuration's whole run: c1 2000, c2 20000, c3 10000, "
                         "c4 5000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--window", type=int, default=500, help="steps per timed window (the run's time profile)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the timed steps one kernel at a time instead of replaying the captured graph "
                         "(for kernel-trace profiling: rocprofv3's tracer crashes on long graph replays)")
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "learned", "surrogate", "cpu-plan", "c3-ref"])
    ap.add_argument("--train-seconds", type=float, default=900.0, help="learned: stop training after this long")
    ap.add_argument("--train-iters", type=int, default=1500, help="learned: Adam iterations (minibatches of 4)")
    ap.add_argument("--train-members", type=int, default=8, help="learned: sweep members (dx = 0.002 m) recorded")
    ap.add_argument("--seed", type=int, default=0, help="learned: training seed (snapshot order, initial weights)")
    ap.add_argument("--k", type=int, default=10, help="learned: plain steps per network step (a configuration of "
                    "its own: the default line evaluates the committed k = 10 weights)")
    ap.add_argument("--save-weights", default="", help="learned: weights file (default gpurun_out/surrogate_seedS.npz)")
    ap.add_argument("--learned-t", type=float, default=2.0, help="simulated time of the learned-vs-plain runs "
                    "(0: none)")
    ap.add_argument("--weights", default=str(WEIGHTS), help="trained surrogate evaluated on the default line")
    ap.add_argument("--learned-max-skip", type=int, default=LEARNED_MAX_SKIP,
                    help="learned rollout: longest wait (intervals) before the next network attempt after "
                         "rejections (0: attempt every interval)")
    ap.add_argument("--no-learned", action="store_true", help="leave the learned-vs-plain evaluation off the line")
    ap.add_argument("--timing-steps", type=int, default=1,
                    help="eager steps with per-kernel HIP events after every timed window")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nbr-block", type=int, default=0, help="fixed launch shape (default: the context's own)")
    ap.add_argument("--nbr-unroll", type=int, default=0)
    ap.add_argument("--xcd-remap", type=int, default=-1)
    ap.add_argument("--nbr-lds", type=int, default=-1)
    ap.add_argument("--nbr-compact", type=int, default=-1)
    ap.add_argument("--nbr-prefetch", type=int, default=-1)
    ap.add_argument("--nbr-list", type=int, default=-1)
    ap.add_argument("--cell-div", type=int, default=0, help="1: the case's 2h grid, 2: half the cell edge "
                    "(with --nbr-lds 4)")
    ap.add_argument("--output-every", type=int, default=0, metavar="M",
                    help="also time VTK frame output every M steps: asynchronous writer vs download-and-write")
    ap.add_argument("--output-frames", type=int, default=4)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a device-to-device copy of MB MiB 5 times (scales HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.config == "learned":
        return learned_bench(args)
    if args.config == "surrogate":
        return surrogate_bench(args)
    if args.config == "cpu-plan":
        return cpu_plan(args)
    if args.config == "c3-ref":
        return reference_surface_bench(args)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): there is no multi-GPU path; "
              "the slab decomposition it would use is described in the design notes", file=sys.stderr)
        return 2

    from tmp_sph_amd import sph
    t0 = time.time()
    desc, cases = _workload(args.config)
    case = cases[0]
    prm = case.params
    n_total = sum(c.n for c in cases)
    n_fluid = sum(c.n_fluid for c in cases)
    print(f"# workload {args.config}: {len(cases)} case(s), N={n_total} fluid={n_fluid} "
          f"(generated in {time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    sim = sph.Simulation(cases if len(cases) > 1 else case)
    fixed = {k: v for k, v in (("nbr_block", args.nbr_block), ("nbr_unroll", args.nbr_unroll)) if v > 0}
    fixed.update({k: v for k, v in (("xcd_remap", args.xcd_remap), ("nbr_lds", args.nbr_lds),
                                     ("nbr_compact", args.nbr_compact), ("nbr_prefetch", args.nbr_prefetch),
                                     ("nbr_list", args.nbr_list))
                  if v >= 0})
    if args.cell_div > 0:
        fixed["cell_div"] = args.cell_div
    if fixed:
        sim.set_tuning(**fixed)
    tuned = sim.tuning.as_dict()
    sim.capture()
    if args.steps <= 0:
        args.steps = RUN_STEPS[args.config]
    sim.replay(args.warmup)
    sim.synchronize()
    # the timed steps in windows (one synchronisation each): the time profile
    # of the run, from the lattice at rest to the developed flow.  After every
    # window, outside the timed region, `timing_steps` eager steps record the
    # per-kernel device time with HIP events on the simulation stream, so the
    # kernel averages sample the whole run.
    sim.set_tuning(flags=1)
    windows = []
    elapsed = 0.0
    done = 0
    extra = 0
    while done < args.steps:
        k = min(args.window, args.steps - done)
        tw = time.perf_counter()
        if args.eager:
            sim.step(k)  # (with per-kernel events when timing is on; only for profiler runs)
        else:
            sim.replay(k)
        sim.synchronize()
        dt_w = time.perf_counter() - tw
        elapsed += dt_w
        windows.append({"first_step": args.warmup + done + extra, "steps": k, "ms_per_step": 1000.0 * dt_w / k})
        done += k
        print(f"#   window {len(windows)}: {done} of {args.steps} steps, {1000.0 * dt_w / k:.3f} ms/step",
              file=sys.stderr, flush=True)
        sim.step(args.timing_steps)
        sim.synchronize()
        extra += args.timing_steps
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n_fluid * args.steps / elapsed
    print(f"# timed {args.steps} steps: {ms_per_step:.3f} ms/step (first window {windows[0]['ms_per_step']:.3f}, "
          f"last {windows[-1]['ms_per_step']:.3f})", file=sys.stderr, flush=True)

    output = output_bench(sim, args.output_every, args.output_frames, ms_per_step) if args.output_every > 0 else None

    kt = sim.kernel_times()
    stats = sim.pair_stats()
    front, height = sim.probe(band=0.02)
    sim.close()
    dom = max(("nsA", "nsB", "nbrlist"), key=lambda k: kt.get(k, 0.0))
    us = kt[dom]
    byts = _algorithmic_bytes(dom, prm.dim, n_total, n_fluid)
    achieved = byts / (us * 1e-6) / 1e9
    # counted flops of the pair code (an fma counts two): the cut-off test 8 per
    # candidate in 3-D (5 in 2-D); an in-support pair 40 in kernel B (35 in
    # 2-D), 17 in kernel A's fluid body (14 in 2-D)
    d3 = prm.dim == 3
    ft = 8.0 if d3 else 5.0
    flops = (ft * stats["candB"] + (40.0 if d3 else 35.0) * stats["hitsB"]) if dom == "nsB" else \
        (ft * stats["candA"] + (17.0 if d3 else 14.0) * stats["hitsA"])
    valu = flops / (us * 1e-6) / 1e12
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{dom}_{args.config}.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
    out = {
        "metric": "particle-steps/s @2M (plain) + learned vs plain speedup at <=2% front error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 positions)",
        "data": "synthetic dam-break lattice released from rest (hydrostatic pressure); the timed steps cover "
                "the collapse and the surge along the tank",
        "config": {"workload": desc, "instances": len(cases), "particles_total": n_total,
                   "particles_fluid": n_fluid, "dt": prm.timeStep if len(cases) == 1 else "per instance",
                   "kernel_h_over_dx": 1.3, "eos": "tait7" if prm.eos == 1 else "linear",
                   "parallelism": "single GPU", "tuning": tuned},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_us": us,
                     "algorithmic_bytes_per_launch": byts,
                     "limiter": "VALU instruction issue with divergent lanes (the pair body runs whenever any "
                                "lane of a wave accepts a candidate), not HBM",
                     "valu": {"achieved_tflops": valu, "peak_tflops": FP32_VALU_PEAK_TFLOPS,
                              "frac": valu / FP32_VALU_PEAK_TFLOPS,
                              "frac_of_non_packed_peak": valu / (FP32_VALU_PEAK_TFLOPS / 2),
                              "flops_per_launch": flops}},
        "windows": windows,
        "t_sim_end_s": (args.warmup + args.steps + extra) * prm.timeStep if len(cases) == 1 else None,
        "kernels_us": kt,
        "kernels_sampled": f"{extra} eager steps, {args.timing_steps} after each timed window (average over the run)",
        "pair_stats": stats,
        "probe_after_run": {"front_m": front, "column_height_m": height},
    }
    if output:
        out["output"] = output
    if not args.no_cpu_baseline and len(cases) == 1:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    if not args.no_learned and args.learned_t > 0 and args.config == "c3":
        if Path(args.weights).exists():
            out["learned"] = learned_eval(Path(args.weights), args.learned_t, max_skip=args.learned_max_skip)
        else:
            out["learned"] = {"skipped": f"no trained weights at {args.weights}"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
