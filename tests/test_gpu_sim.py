"""GPU parity of the fast path (sph_sim_*: sorted SoA state, FP64 positions,
hipGraph replay) against the oracle's FP32-fields / FP64-positions mode."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import advanced, case2d, case3d, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _worst(case, name, got, want, mag, slack=0.0):
    g = np.asarray(got, np.float64).reshape(len(case.type), -1)
    w = np.asarray(want, np.float64).reshape(len(case.type), -1)
    m = np.asarray(mag, np.float64).reshape(len(case.type), -1)
    sl = np.broadcast_to(np.asarray(slack, np.float64).reshape(-1, 1) if np.ndim(slack) else slack, g.shape)
    fl = case.type == 1
    r = np.where(fl[:, None], np.maximum(np.abs(g - w) - sl, 0.0) / np.maximum(np.maximum(np.abs(w), m), 1e-300), 0.0)
    i, c = np.unravel_index(np.argmax(r), r.shape)
    print(f"worst {name}: particle {i} comp {c} pos {case.pos[i]} got {g[i, c]:.9g} want {w[i, c]:.9g} "
          f"mag {m[i, c]:.6g} rel {r[i, c]:.3e}")


def _check_rates(case, got, want):
    prm = case.params
    np.testing.assert_array_equal(got["keys"], want["keys"])
    np.testing.assert_array_equal(got["perm"], want["perm"])
    fl = case.type == 1
    wl = ~fl
    _worst(case, "drho", got["drho"], want["drho"], want["mag_drho"], want["sens_drho"])
    _worst(case, "acc", got["acc"], want["acc"], want["mag_acc"])
    assert rel_err(got["drho"][fl], want["drho"][fl], want["mag_drho"][fl], want["sens_drho"][fl]) < TOL
    assert rel_err(got["acc"][fl], want["acc"][fl], want["mag_acc"][fl]) < TOL
    assert rel_err(got["dofv"], want["dofv"], want["mag_dofv"]) < TOL
    pscale = prm.restDensity * abs(prm.gravity) * prm.h
    assert rel_err(got["p"][wl], want["p"][wl], pscale) < TOL
    assert rel_err(got["rho"][wl], want["rho"][wl], prm.restDensity) < TOL
    np.testing.assert_array_equal(got["p"][fl], want["p"][fl])
    np.testing.assert_array_equal(got["rho"][fl], want["rho"][fl])


@pytest.mark.parametrize("which", ["2d_rest", "2d_flow", "3d_flow"])
def test_eval_rates_match_oracle(gpu, orc, which):
    if which == "2d_rest":
        case = case2d(0.01)
    elif which == "2d_flow":
        case = advanced(orc, case2d(0.02), 200)
    else:
        case = advanced(orc, case3d(0.04), 50)
    with gpu.Simulation(case) as sim:
        got = sim.eval_rates()
    want = orc.eval_rates(case.params, case, mode=1)
    _check_rates(case, got, want)


@pytest.mark.parametrize("which", ["2d_flow", "3d_flow"])
@pytest.mark.parametrize("cell_div", [1, 2])
def test_fine_rows_match_oracle(gpu, orc, which, cell_div):
    """Fine rows (nbr_lds 4) on the instances' grid and on the grid of half
    the cell edge: keys and order bit-exact against the oracle run on that
    grid, rates within 1e-5 of it (the oracle visits 2K+1 cells per axis in
    the same row order)."""
    case = advanced(orc, case2d(0.02), 200) if which == "2d_flow" else advanced(orc, case3d(0.04), 50)
    with gpu.Simulation(case) as sim:
        sim.set_tuning(nbr_lds=4, cell_div=cell_div)
        got = sim.eval_rates()
    grid = case.copy()
    grid.params = gpu.subdivide_params(case.params, cell_div)
    want = orc.eval_rates(grid.params, grid, mode=1)
    _check_rates(grid, got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_fine_rows_visit_the_same_pairs(gpu, dim):
    """The in-support pairs the fine rows visit on the halved grid are those
    of the 2h grid (statistics of both visits), with fewer candidates."""
    case = case2d(0.01) if dim == 2 else case3d(0.03)
    with gpu.Simulation(case) as sim:
        sim.step(3)
        coarse = sim.pair_stats()
        sim.set_tuning(nbr_lds=4, cell_div=1)
        fine1 = sim.pair_stats()
        sim.set_tuning(nbr_lds=4, cell_div=2)
        fine2 = sim.pair_stats()
    print(dim, coarse, fine1, fine2)
    for k in ("hitsA", "hitsB"):
        assert fine1[k] == coarse[k] and fine2[k] == coarse[k], k
    assert fine2["candB"] < coarse["candB"]


@pytest.mark.parametrize("dim,steps", [(2, 20), (3, 8)])
def test_steps_match_oracle(gpu, orc, dim, steps):
    case = advanced(orc, case2d(0.02) if dim == 2 else case3d(0.04), 30)
    prm = case.params
    with gpu.Simulation(case) as sim:
        sim.step(steps)
        got = sim.state()
        got64 = sim.pos64()
    want, _, _, _ = orc.run(prm, case, steps, mode=1)
    fl = case.type == 1
    # rounding-level differences in the rates accumulate over the steps
    dv = np.abs(got["vel"][fl] - want["vel"][fl]).max()
    vscale = np.abs(want["vel"][fl]).max()
    assert dv <= 1e-4 * vscale + 1e-6, dv
    dx = np.abs(got64[fl] - want["pos64"][fl]).max()
    assert dx <= 1e-4 * vscale * steps * prm.timeStep + 1e-9, dx
    assert rel_err(got["rho"][fl], want["rho"][fl], prm.restDensity) < 1e-5
    np.testing.assert_array_equal(got["type"], want["type"])
    # walls never move
    np.testing.assert_array_equal(got["pos"][~fl], case.pos[~fl])


def test_graph_replay_matches_eager(gpu):
    case = case2d(0.02)
    with gpu.Simulation(case) as a, gpu.Simulation(case) as b:
        a.step(6)
        b.capture()
        b.replay(6)
        b.synchronize()
        sa, sb = a.state(), b.state()
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
        np.testing.assert_array_equal(a.pos64(), b.pos64())
        assert b.steps_done == 6


def test_block_size_does_not_change_results(gpu):
    case = case2d(0.02)
    out = []
    for blk in (64, 128, 256):
        with gpu.Simulation(case) as s:
            s.set_tuning(nbr_block=blk)
            s.step(4)
            out.append(s.pos64())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_probe_matches_numpy(gpu, orc):
    case = advanced(orc, case2d(0.02), 300)
    with gpu.Simulation(case) as sim:
        f, h = sim.probe(band=0.05)
        p64 = sim.pos64()
    fl = case.type == 1
    assert f == pytest.approx(p64[fl, 0].max(), abs=0)
    sel = fl & (p64[:, 0] <= 0.05)
    assert h == pytest.approx(p64[sel, 2].max(), abs=0)


def test_probe_with_a_domain_at_negative_coordinates(gpu):
    """The probe's maxima hold for any sign: the whole case (tank, grid,
    escape box) shifted to negative x and z."""
    case = case2d(0.02)
    shift = np.array([-7.0, 0.0, -3.0], np.float32)
    case.pos = (case.pos + shift).astype(np.float32)
    prm = case.params
    prm.gridxmin += float(shift[0])
    prm.gridzmin += float(shift[2])
    for k in (0, 2):
        prm.domain_min[k] += float(shift[k])
        prm.domain_max[k] += float(shift[k])
    prm.park_pos[0] += float(shift[0])
    prm.park_pos[2] += float(shift[2])
    band = float(shift[0]) + 0.05
    with gpu.Simulation(case) as sim:
        sim.step(3)
        f, h = sim.probe(band=band)
        p64 = sim.pos64()
    fl = case.type == 1
    assert f == p64[fl, 0].max() and f < 0
    sel = fl & (p64[:, 0] <= band)
    assert h == p64[sel, 2].max() and h < 0


def test_escaped_particle_is_parked(gpu):
    case = case2d(0.02)
    fl = np.flatnonzero(case.type == 1)
    case.pos[fl[-1]] = (case.params.domain_max[0] + 0.005, 0.0, 0.1)
    with gpu.Simulation(case) as sim:
        sim.step(1)
        st = sim.state()
    np.testing.assert_array_equal(st["pos"][fl[-1]], np.asarray(case.params.park_pos, np.float32))
    assert np.all(st["vel"][fl[-1]] == 0) and st["p"][fl[-1]] == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_lone_fluid_particle_meets_only_itself(gpu, dim):
    """A fluid particle with nothing else in reach (in the air, away from the
    column and the walls) pairs only with itself in the neighbour loops, and
    that pair's terms are exactly zero (dx = dy = dz = 0, dvdx = 0, equal
    densities): drho = dofv = 0 and acc = (0, 0, g) bit for bit, for the
    default row-staged loop, the plain global loop and the fine rows."""
    case = case2d(0.02) if dim == 2 else case3d(0.04)
    k = int(np.flatnonzero(case.type == 1)[-1])
    case.pos[k] = (3.0, 0.0, 0.7) if dim == 2 else (0.6, 0.2, 0.35)
    case.vel[k] = (0.3, 0.1 if dim == 3 else 0.0, -0.2)
    case.rho[k] = case.params.restDensity * 1.002
    case.p[k] = 500.0
    reach = np.sqrt(np.float64(case.params.rr_cut))
    others = np.delete(case.pos.astype(np.float64), k, axis=0)
    assert np.min(np.linalg.norm(others - case.pos[k].astype(np.float64), axis=1)) > 1.5 * reach
    shapes = ({}, dict(nbr_lds=0, nbr_unroll=2, nbr_block=256, nbr_prefetch=0), dict(nbr_lds=4, cell_div=1))
    with gpu.Simulation(case) as sim:
        for shape in shapes:
            sim.set_tuning(**shape)
            r = sim.eval_rates()
            assert r["drho"][k] == 0.0 and r["dofv"][k] == 0.0, shape
            np.testing.assert_array_equal(r["acc"][k], np.float32([0.0, 0.0, case.params.gravity]), err_msg=str(shape))


def _gpu_flow_state(gpu, case, steps):
    """The case advanced `steps` steps by the GPU (graph replay), returned as a
    new case in original particle order: a developed flow state at full size
    that the oracle can evaluate once."""
    with gpu.Simulation(case) as sim:
        sim.capture()
        sim.replay(steps)
        sim.synchronize()
        st = sim.state()
    out = case.copy()
    out.pos, out.vel, out.rho, out.p, out.type = st["pos"], st["vel"], st["rho"], st["p"], st["type"]
    return out


@pytest.mark.parametrize("steps", [0, 2000])
def test_full_size_c2(gpu, orc, steps):
    """2-D dam break at dx = 0.001 (500k fluid + 18k wall), at rest and after
    2000 steps: hash, sorted order bit-exact and rates within 1e-5 at the
    full size."""
    case = case2d(0.001)
    if steps:
        case = _gpu_flow_state(gpu, case, steps)
        assert np.abs(case.vel[case.type == 1]).max() > 0.01  # the column has started to fall
    with gpu.Simulation(case) as sim:
        got = sim.eval_rates()
    want = orc.eval_rates(case.params, case, mode=1)
    _check_rates(case, got, want)


@pytest.mark.parametrize("steps", [0, 1000])
def test_full_size_c3(gpu, orc, steps):
    """3-D dam break at dx = 0.005 (2.0M fluid + 1.41M wall), at rest and after
    1000 steps: hash and sorted order bit-exact against the oracle, rates
    within 1e-5, and the default launch shape (row staging with its fallback
    waves, 23-bit keys) bit-identical to the plain global loop."""
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    case = make_case(dambreak_cfg(dim=3, dx=0.005, tank=(3.2, 1.0, 1.0), column=(1.0, 0.5, 0.5)))
    assert case.n_fluid == 2_000_000
    if steps:
        case = _gpu_flow_state(gpu, case, steps)
        assert np.abs(case.vel[case.type == 1]).max() > 0.05
    with gpu.Simulation(case) as sim:
        got = sim.eval_rates()
        sim.set_tuning(nbr_lds=0, nbr_compact=0, nbr_list=0, nbr_prefetch=0, nbr_block=256)
        plain = sim.eval_rates()
    for k in ("drho", "acc", "dofv", "p", "rho", "keys", "perm"):
        np.testing.assert_array_equal(plain[k], got[k], err_msg=k)
    want = orc.eval_rates(case.params, case, mode=1)
    _check_rates(case, got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_gauges_and_pressure_probes_match_numpy(gpu, dim):
    """The reference's measuring probes on the device: height gauges (largest
    fluid z in an x (and y) band under a cap) and point pressures by id."""
    case = case2d(0.01) if dim == 2 else case3d(0.03)
    with gpu.Simulation(case) as sim:
        sim.step(30)
        st = sim.state()
        dx = float(case.params.particleR)
        hw = 1.5 * dx
        gx = np.array([0.05, 0.2, 0.5, 3.9], np.float32)
        gy = np.full(4, 0.5 * float(case.params.domain_max[1]) if dim == 3 else 0.0, np.float32)
        zc = np.array([10.0, 0.2, 10.0, 10.0], np.float32)
        got = sim.gauges(gx, gy, zc, hw)
        rng = np.random.default_rng(0)
        ids = rng.choice(case.n, 8, replace=False)
        p = sim.pressure_at(ids)
    pos, fl = st["pos"], st["type"] == 1
    for g in range(4):
        sel = fl & (np.abs(pos[:, 0] - gx[g]) < np.float32(hw)) & (pos[:, 2] < zc[g])
        if dim == 3:
            sel &= np.abs(pos[:, 1] - gy[g]) < np.float32(hw)
        want = max(float(pos[sel, 2].max()), 0.0) if sel.any() else 0.0
        assert got[g] == np.float32(want), (g, got[g], want)
    assert got[0] > 0 and got[3] == 0  # a band through the column, one beyond the water
    np.testing.assert_array_equal(p, st["p"][ids])


@pytest.mark.parametrize("dim,steps,every", [(2, 2400, 100), (3, 300, 50)])
def test_front_and_height_histories_match_fp64_host_run(gpu, orc, dim, steps, every):
    """Surge-front and water-column-height histories of the GPU run against
    the oracle's all-FP64 run of the same case (2-D: t = 0.4 s, the front
    travels from 1.0 to 2.0 m).  The host's own FP32 and FP64 runs agree to
    1e-6 m here, so 1e-4 m (a fifth of a percent of the spacing) leaves room
    only for rounding-level differences, not for a wrong term."""
    case = case2d(0.02) if dim == 2 else case3d(0.04)
    band = 2.0 * case.params.particleR
    fr, ht = [], []
    with gpu.Simulation(case) as sim:
        sim.capture()
        for _ in range(steps // every):
            sim.replay(every)
            f, h = sim.probe(band)
            fr.append(f)
            ht.append(h)
    _, f64, h64, _ = orc.run(case.params, case, steps, mode=2, band=band, probe_every=every)
    fr, ht = np.array(fr), np.array(ht)
    print(f"dim {dim}: front {fr[0]:.4f} -> {fr[-1]:.4f} m, max |GPU - FP64| front {np.abs(fr - f64).max():.2e}, "
          f"height {np.abs(ht - h64).max():.2e}")
    assert fr[-1] > fr[0] + (0.5 if dim == 2 else 0.05)  # the surge has moved
    np.testing.assert_allclose(fr, f64, rtol=0, atol=1e-4)
    np.testing.assert_allclose(ht, h64, rtol=0, atol=1e-4)
