"""GPU parity of the reference call surface's physics: SPH_NS_simpleversion
(kernels A and B, opt_src_v1/equation/ns.cu:5-230) within 1e-5 of the oracle
(magnitude-normalised), PC_prediction / PC_correction bit-exact
(opt_src_v0/sph/sph_run.cu:388-484), and one whole reference-order step."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import advanced, case2d, case3d, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative, per field, against max(|oracle|, magnitude sum of its terms)


def _sorted_state(orc, case):
    prm = case.params
    h, i = orc.calcHash(prm, case.pos)
    hs, is_ = orc.sortParticles(prm, h, i)
    st = orc.reorder(prm, case.pos, case.vel, case.rho, case.p, case.type, hs, is_)
    st["hash"], st["idx"] = hs, is_
    return st


def _gpu_ns(sph, prm, st, mode=0):
    n = st["pos"].shape[0]
    D = sph.DeviceArray.from_host
    d = {k: D(st[k]) for k in ("pos", "vel", "rho", "p", "type", "cellStart", "cellEnd", "hash")}
    drho = sph.DeviceArray((n,), np.float32)
    acc = sph.DeviceArray((n, 3), np.float32)
    dofv = sph.DeviceArray((n,), np.float32)
    sph.set_params(prm)
    sph.set_ns_mode(mode)
    try:
        sph.SPH_NS_simpleversion(d["pos"], d["vel"], d["rho"], d["p"], d["type"], drho, acc, d["cellStart"],
                                 d["cellEnd"], n, d["hash"], 0, dofv)
    finally:
        sph.set_ns_mode(0)
    return {"drho": drho.numpy(), "acc": acc.numpy(), "dofv": dofv.numpy(), "rho": d["rho"].numpy(),
            "p": d["p"].numpy()}


def _pressure_sensitivity(prm, st, rho):
    """Per particle, sum over its in-support neighbours j of
    m |F(r_ij)| |r_ij| / (rho_i rho_j): the acceleration a pressure change of 1
    on every neighbour can cause through the pressure term of kernel B
    (F = -5 adh (1 - q/2)^3 / h^2, the kernel gradient factor)."""
    from scipy.spatial import cKDTree
    pos = st["pos"].astype(np.float64)
    pairs = cKDTree(pos).query_pairs(float(np.sqrt(prm.rr_cut)) * 1.0001, output_type="ndarray")
    d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
    r = np.sqrt((d * d).sum(1))
    q = r / prm.h
    f = np.where(q <= 2.0, 5.0 * prm.adh * (1.0 - q / 2.0) ** 3 / (prm.h * prm.h), 0.0)
    term = prm.particleMass * f * r / (np.abs(rho[pairs[:, 0]]) * np.abs(rho[pairs[:, 1]]))
    out = np.zeros(pos.shape[0])
    np.add.at(out, pairs[:, 0], term)
    np.add.at(out, pairs[:, 1], term)
    return out


def _check_ns(prm, got, want, fluid, wall, st=None):
    """Rates within TOL of the oracle, magnitude-normalised.  With `st` (the
    fast loops, whose wall pressures carry their own 1e-5 x rho0 g h
    contract) a wall target's acceleration is also allowed what that
    pressure tolerance on its neighbours can move it: near the free surface a
    wall's acceleration is the tiny residual of a barely wet neighbour's
    extrapolated pressure, itself a small difference of large terms."""
    assert rel_err(got["drho"][fluid], want["drho"][fluid], want["mag_drho"][fluid], want["sens_drho"][fluid]) < TOL
    assert rel_err(got["acc"][fluid], want["acc"][fluid], want["mag_acc"][fluid]) < TOL
    assert rel_err(got["dofv"], want["dofv"], want["mag_dofv"]) < TOL
    pscale = prm.restDensity * abs(prm.gravity) * prm.h
    assert rel_err(got["p"][wall], want["p"][wall], pscale) < TOL
    assert rel_err(got["rho"][wall], want["rho"][wall], prm.restDensity) < TOL
    # walls: the reference computes kernel B for them too
    assert rel_err(got["drho"][wall], want["drho"][wall], want["mag_drho"][wall], want["sens_drho"][wall]) < TOL
    scale = want["mag_acc"][wall]
    if st is not None:
        sens = pscale * _pressure_sensitivity(prm, st, want["rho"])[wall]
        scale = np.maximum(scale, sens[:, None])
    assert rel_err(got["acc"][wall], want["acc"][wall], scale) < TOL


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("which", ["2d_tait", "2d_linear", "3d_tait", "ref"])
def test_sph_ns_matches_oracle(gpu, orc, which, mode):
    """SPH_NS_simpleversion on the fast loops (mode 0, the default) and in the
    reference's operation order (mode 1) against the oracle, every particle
    (walls' rates included); mode 1 is bit-exact with the linear EOS."""
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    if which == "2d_tait":
        case = advanced(orc, case2d(0.02), 150, mode=0)
    elif which == "2d_linear":
        case = advanced(orc, case2d(0.02, eos=0, c0=20.0), 150, mode=0)
    elif which == "3d_tait":
        case = advanced(orc, case3d(0.04), 40, mode=0)
    else:
        # the reference's parameter recipe (linear EOS, h = 1.5 dx, delta 0.2, afa 0.05)
        c = make_case(dambreak_cfg(dim=3, dx=0.02, tank=(0.6, 0.3, 0.4), column=(0.2, 0.3, 0.2), eos=0))
        from tmp_sph_amd.sph import reference_params
        prm = reference_params(dx=0.02, num_particles=c.n, num_fluid=c.n_fluid)
        c.params = prm
        case = advanced(orc, c, 20, mode=0)
    prm = case.params
    st = _sorted_state(orc, case)
    want = orc.SPH_NS(prm, st["pos"], st["vel"], st["rho"], st["p"], st["type"], st["cellStart"], st["cellEnd"],
                      st["hash"])
    got = _gpu_ns(gpu, prm, st, mode)
    fluid = st["type"] == 1
    _check_ns(prm, got, want, fluid, ~fluid, st if mode == 0 else None)
    assert np.abs(want["acc"][fluid]).max() > 0  # a non-trivial flow state
    if prm.eos == 0 and mode == 1:
        # linear EOS: same operations in the same order with no contraction,
        # so the reference surface reproduces the oracle bit for bit
        for k in ("drho", "acc", "dofv", "p", "rho"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_pc_bit_exact(gpu, orc):
    sph = gpu
    case = advanced(orc, case2d(0.02), 60, mode=0)
    prm = case.params
    st = _sorted_state(orc, case)
    n = case.n
    rng = np.random.default_rng(5)
    drho = (rng.standard_normal(n) * 50).astype(np.float32)
    acc = (rng.standard_normal((n, 3)) * 20).astype(np.float32)
    # a few fluid particles outside the escape box must be parked
    fl = np.flatnonzero(st["type"] == 1)
    out_idx = fl[:5]
    st["pos"][out_idx, 0] = np.float32(prm.domain_max[0] + 0.01)
    tpos = np.zeros((n, 3), np.float32)
    tvel = np.zeros((n, 3), np.float32)
    trho = np.zeros(n, np.float32)
    want1 = orc.PC(prm, True, st["pos"], st["vel"], st["rho"], st["p"], drho, acc, None, tpos, tvel, trho,
                   st["type"])
    D = sph.DeviceArray.from_host
    sph.set_params(prm)
    d = {k: D(st[k]) for k in ("pos", "vel", "rho", "p", "type", "hash", "idx")}
    ddr, dacc = D(drho), D(acc)
    dt_pos, dt_vel, dt_rho = D(tpos), D(tvel), D(trho)
    sph.PC_prediction(d["pos"], d["vel"], d["rho"], d["p"], ddr, dacc, d["hash"], d["idx"], dt_pos, dt_vel, dt_rho,
                      n, d["type"])
    for k, dk in (("pos", "pos"), ("vel", "vel"), ("rho", "rho"), ("p", "p"), ("type", "type")):
        np.testing.assert_array_equal(d[dk].numpy(), want1[k], err_msg=f"predictor {k}")
    fl1 = want1["type"] == 1
    np.testing.assert_array_equal(dt_pos.numpy()[fl1], want1["tpos"][fl1])
    np.testing.assert_array_equal(dt_rho.numpy()[fl1], want1["trho"][fl1])
    assert np.all(want1["pos"][out_idx] == np.asarray(prm.park_pos, np.float32))
    # corrector on a shuffled order: slot i holds the particle predicted at slot perm[i]
    perm = rng.permutation(n).astype(np.int32)
    s2 = {k: want1[k][perm] for k in ("pos", "vel", "rho", "p", "type")}
    want2 = orc.PC(prm, False, s2["pos"], s2["vel"], s2["rho"], s2["p"], drho, acc, perm, want1["tpos"],
                   want1["tvel"], want1["trho"], s2["type"])
    e = {k: D(s2[k]) for k in s2}
    dperm = D(perm)
    sph.PC_correction(e["pos"], e["vel"], e["rho"], e["p"], ddr, dacc, d["hash"], dperm, dt_pos, dt_vel, dt_rho, n,
                      e["type"])
    for k in ("pos", "vel", "rho", "p", "type"):
        np.testing.assert_array_equal(e[k].numpy(), want2[k], err_msg=f"corrector {k}")


def _ref_surface_run(sph, case, nsteps, mode):
    """`nsteps` full steps through the reference call surface in the order of
    opt_src_v0/sph/sph_run.cu:269-309 (per half-step calcHash, sortParticles,
    reorder with the permutation table, SPH_NS_simpleversion, PC); returns
    the state in original particle order."""
    prm = case.params
    n = case.n
    sph.set_params(prm)
    sph.set_ns_mode(mode)
    sph.reserve_scratch(n)  # the calls below then allocate nothing
    D, E = sph.DeviceArray.from_host, sph.DeviceArray
    S = [{"pos": D(case.pos), "vel": D(case.vel), "rho": D(case.rho), "p": D(case.p), "type": D(case.type)},
         {"pos": E((n, 3), np.float32), "vel": E((n, 3), np.float32), "rho": E((n,), np.float32),
          "p": E((n,), np.float32), "type": E((n,), np.int32)}]
    tabs = [D(np.zeros(n, np.int32)), D(np.zeros(n, np.int32))]
    h, idx = E((n,), np.int32), E((n,), np.int32)
    cs, ce = E((prm.numCells,), np.int32), E((prm.numCells,), np.int32)
    drho, acc, dofv = E((n,), np.float32), E((n, 3), np.float32), E((n,), np.float32)
    tpos, tvel, trho = E((n, 3), np.float32), E((n, 3), np.float32), E((n,), np.float32)
    try:
        for _ in range(nsteps):
            for a, b, pred in ((0, 1, True), (1, 0, False)):
                A, B = S[a], S[b]
                sph.calcHash(A["pos"], h, idx, n)
                sph.sortParticles(h, idx, n)
                sph.reorder(A["pos"], A["vel"], A["rho"], A["p"], A["type"], None, B["pos"], B["vel"], B["rho"],
                            B["p"], B["type"], None, h, idx, cs, ce, n, tabs[a], tabs[b])
                sph.SPH_NS_simpleversion(B["pos"], B["vel"], B["rho"], B["p"], B["type"], drho, acc, cs, ce, n, h,
                                         0, dofv)
                f = sph.PC_prediction if pred else sph.PC_correction
                f(B["pos"], B["vel"], B["rho"], B["p"], drho, acc, h, idx, tpos, tvel, trho, n, B["type"])
    finally:
        sph.set_ns_mode(0)
    ids = np.arange(n, dtype=np.int32) + tabs[1].numpy()  # i + table[i] = original id (the second half wrote tabs[1])
    assert np.array_equal(np.sort(ids), np.arange(n))
    got = {k: np.empty_like(v.numpy()) for k, v in S[0].items()}
    for k, v in S[0].items():
        got[k][ids] = v.numpy()
    return got


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_order_step(gpu, orc, mode):
    """One step through the reference call surface against the oracle's FP32
    step, with SPH_NS_simpleversion on the fast loops (mode 0) and in
    reference order (mode 1)."""
    case = advanced(orc, case2d(0.02), 100, mode=0)
    prm = case.params
    got = _ref_surface_run(gpu, case, 1, mode)
    want, _, _, _ = orc.run(prm, case, 1, mode=0)
    mags = orc.eval_rates(prm, case, mode=0)
    dt = prm.timeStep
    fl = case.type == 1
    vtol = TOL * mags["mag_acc"][fl].max() * dt
    assert np.max(np.abs(got["vel"][fl] - want["vel"][fl])) <= vtol + 1e-6 * np.abs(want["vel"][fl]).max()
    assert np.max(np.abs(got["pos"][fl] - want["pos"][fl])) <= vtol * dt + 4e-7
    rtol = TOL * mags["mag_drho"][fl].max() * dt
    assert np.max(np.abs(got["rho"][fl] - want["rho"][fl])) <= rtol + 1e-6 * prm.restDensity
    np.testing.assert_array_equal(got["type"], want["type"])


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_surface_run_follows_the_host_run(gpu, orc, mode):
    """600 steps (t = 0.1 s, the column collapsing) of the 2-D dam break at
    dx = 0.02 m through the reference call surface: the surge front and
    every fluid particle's position stay within 1e-4 m of the oracle's FP32
    run from the same state (the step-to-step differences are rounding-level,
    1e-5 of the rates)."""
    case = case2d(0.02)
    nsteps = 600
    got = _ref_surface_run(gpu, case, nsteps, mode)
    want, _, _, _ = orc.run(case.params, case, nsteps, mode=0)
    fl = want["type"] == 1
    np.testing.assert_array_equal(got["type"], want["type"])
    front_got, front_want = got["pos"][fl, 0].max(), want["pos"][fl, 0].max()
    assert front_want > 1.05  # the surge has left the column (1.0 m)
    assert abs(front_got - front_want) <= 1e-4, (front_got, front_want)
    assert np.abs(got["pos"][fl] - want["pos"][fl]).max() <= 1e-4
