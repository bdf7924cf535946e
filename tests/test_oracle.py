"""Pins of the CPU oracle (tests/ only).  The reference ships no tests, golden
vectors or fixtures, so every check here comes from a library routine, a
closed form, an invariant or a brute-force count, chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails one of them."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import advanced, case2d, case3d, rel_err


def _lattice_case(dim, dx=0.01, side=(0.4, 0.4, 0.4), origin=(1.0, 0.3, 0.3), p_fn=None, rho_fn=None,
                  v_fn=None, eos=0, c0=20.0, h_over_dx=2.0):
    """A fluid-only lattice block (no walls) inside a large grid.  h = 2 dx
    keeps lattice sums within a fraction of a percent of the integrals."""
    from tmp_sph_amd import sph
    cfg = sph.dambreak_cfg(dim=dim, dx=dx, tank=(2.0, 1.0, 1.0), column=(0.1, 0.1, 0.1), eos=eos, c0=c0,
                           h_over_dx=h_over_dx)
    base = sph.make_case(cfg)
    n = [int(round(s / dx)) for s in side]
    if dim == 2:
        n[1] = 1
    ii = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    pos = (np.asarray(origin) + (ii + 0.5) * dx).astype(np.float32)
    if dim == 2:
        pos[:, 1] = 0.0
    N = pos.shape[0]
    rho = np.full(N, 1000.0, np.float32) if rho_fn is None else rho_fn(pos).astype(np.float32)
    p = np.zeros(N, np.float32) if p_fn is None else p_fn(pos).astype(np.float32)
    vel = np.zeros((N, 3), np.float32) if v_fn is None else v_fn(pos).astype(np.float32)
    c = base.copy()
    c.pos, c.vel, c.rho, c.p, c.type = pos, vel, rho, p, np.ones(N, np.int32)
    c.n_fluid = N
    c.params.domain_min[:] = (-1.0, -1.0, -1.0)
    c.params.domain_max[:] = (5.0, 5.0, 5.0)
    return c


def _interior(case, margin):
    lo = case.pos.min(0) + margin
    hi = case.pos.max(0) - margin
    m = np.all((case.pos >= lo) & (case.pos <= hi), axis=1) if case.params.dim == 3 else \
        (case.pos[:, 0] >= lo[0]) & (case.pos[:, 0] <= hi[0]) & (case.pos[:, 2] >= lo[2]) & (case.pos[:, 2] <= hi[2])
    assert m.sum() > 10
    return m


# ---------------------------------------------------------------- mesh
def test_hash_on_cell_faces_and_clamping(orc):
    from tmp_sph_amd import sph
    prm = sph.reference_params()
    s = np.float32(prm.gridSize)
    k = np.arange(0, 40)
    # positions built in float32 exactly as the kernel subtracts: min + k*size
    x = (np.float32(prm.gridxmin) + k.astype(np.float32) * s).astype(np.float32)
    pos = np.stack([x, np.full_like(x, 0.5), np.full_like(x, 0.5)], 1)
    h, idx = orc.calcHash(prm, pos)
    cx = h % prm.gridxdim
    want = np.floor((x - np.float32(prm.gridxmin)) / s).astype(np.int64)  # IEEE float32 numpy
    np.testing.assert_array_equal(cx, want)
    assert np.all(np.diff(cx) >= 0)
    np.testing.assert_array_equal(idx, np.arange(len(x)))
    # outside the grid: clamped into the boundary cells
    far = np.array([[-5, -5, -5], [50, 50, 50]], np.float32)
    hf, _ = orc.calcHash(prm, far)
    assert hf[0] == 0 and hf[1] == prm.numCells - 1


def test_sort_is_numpy_stable_argsort(orc):
    from tmp_sph_amd import sph
    prm = sph.reference_params()
    rng = np.random.default_rng(0)
    for hi in (3, 100, prm.numCells):
        keys = rng.integers(0, hi, 50_000).astype(np.int32)
        idx = np.arange(keys.size, dtype=np.int32)
        ks, is_ = orc.sortParticles(prm, keys, idx)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(is_, order)
        np.testing.assert_array_equal(ks, keys[order])


def test_cell_ranges_are_searchsorted(orc):
    case = advanced(orc, case2d(0.02), 40, mode=0)
    prm = case.params
    h, i = orc.calcHash(prm, case.pos)
    hs, is_ = orc.sortParticles(prm, h, i)
    table_r = np.zeros(case.n, np.int32)
    st = orc.reorder(prm, case.pos, case.vel, case.rho, case.p, case.type, hs, is_, table_r)
    cells = np.arange(prm.numCells)
    lo = np.searchsorted(hs, cells, "left")
    hi = np.searchsorted(hs, cells, "right")
    nonempty = hi > lo
    np.testing.assert_array_equal(st["cellStart"][nonempty], lo[nonempty])
    np.testing.assert_array_equal(st["cellEnd"][nonempty], hi[nonempty])
    assert np.all(st["cellStart"][~nonempty] == -1)
    # gather and permutation table: i + table_l[i] is the original id
    np.testing.assert_array_equal(st["pos"], case.pos[is_])
    np.testing.assert_array_equal(np.arange(case.n) + st["table_l"], is_)


# -------------------------------------------------------------------- EOS
def test_eos_closed_forms(orc):
    c2 = case2d(0.02)
    prm = c2.params  # Tait, gamma 7
    B = float(prm.tait_B)
    for rho in (1000.0, 1000.5, 1010.0, 1100.0):
        want = B * ((rho / 1000.0) ** 7 - 1.0)
        assert orc.eos_pressure(prm, rho, fp64=True) == pytest.approx(want, rel=1e-12)
        assert orc.eos_density(prm, orc.eos_pressure(prm, rho, True), True) == pytest.approx(rho, rel=1e-12)
    lin = case2d(0.02, eos=0, c0=25.0).params
    assert orc.eos_pressure(lin, 1004.0) == pytest.approx(625.0 * 4.0, rel=1e-6)
    assert orc.eos_density(lin, 2500.0) == pytest.approx(1004.0, rel=1e-7)


# ---------------------------------------------------------------- kernel A
@pytest.mark.parametrize("dim", [2, 3])
def test_wall_pressure_is_hydrostatic_extrapolation(orc, dim):
    """With rho_j = rho0 and p_j = rho0 |g| (H - z_j), kernel A's wall formula
    gives exactly p_w = rho0 |g| (H - z_w): the hydrostatic correction
    cancels the height difference (a wrong sign of g or dz breaks this)."""
    H = 1.2
    c = _lattice_case(dim, dx=0.02, side=(0.3, 0.3, 0.3), p_fn=lambda x: 1000 * 9.81 * (H - x[:, 2]))
    mid = c.pos.mean(0)
    walls = mid + np.array([[0.013, 0.0, -0.017], [-0.021, 0.0, 0.006], [0.004, 0.0, 0.029]], np.float32)
    if dim == 3:
        walls[:, 1] = mid[1] + np.array([0.011, -0.007, 0.0])
    c.pos = np.concatenate([c.pos, walls.astype(np.float32)])
    c.vel = np.concatenate([c.vel, np.zeros((3, 3), np.float32)])
    c.rho = np.concatenate([c.rho, np.full(3, 1000.0, np.float32)])
    c.p = np.concatenate([c.p, np.zeros(3, np.float32)])
    c.type = np.concatenate([c.type, np.zeros(3, np.int32)])
    out = orc.eval_rates(c.params, c, mode=2)
    pw = out["p"][-3:]
    np.testing.assert_allclose(pw, 1000 * 9.81 * (H - walls[:, 2].astype(np.float64)), rtol=2e-6)
    cs = float(c.params.cs)
    np.testing.assert_allclose(out["rho"][-3:], pw / cs ** 2 + 1000.0, rtol=1e-9)  # linear inverse EOS


@pytest.mark.parametrize("dim", [2, 3])
def test_velocity_divergence_of_linear_field(orc, dim):
    """dofv approximates div v: for v = (a x, 0, b z) it is a + b in the interior."""
    a, b = 0.7, -0.3
    c = _lattice_case(dim, dx=0.01, side=(0.2, 0.2, 0.2),
                      v_fn=lambda x: np.stack([a * x[:, 0], 0 * x[:, 0], b * x[:, 2]], 1))
    out = orc.eval_rates(c.params, c, mode=2)
    m = _interior(c, 2.5 * c.params.h)
    np.testing.assert_allclose(out["dofv"][m], a + b, rtol=0.02)


# ---------------------------------------------------------------- kernel B
@pytest.mark.parametrize("dim", [2, 3])
def test_continuity_of_linear_field(orc, dim):
    """drho/dt = -rho div v for a uniform-density lattice with linear v."""
    a, b = 0.5, 0.2
    c = _lattice_case(dim, dx=0.01, side=(0.2, 0.2, 0.2),
                      v_fn=lambda x: np.stack([a * x[:, 0], 0 * x[:, 0], b * x[:, 2]], 1))
    out = orc.eval_rates(c.params, c, mode=2)
    m = _interior(c, 2.5 * c.params.h)
    np.testing.assert_allclose(out["drho"][m], -1000.0 * (a + b), rtol=0.02)


@pytest.mark.parametrize("dim", [2, 3])
def test_delta_diffusion_of_quadratic_density(orc, dim):
    """At rest with rho = rho0 + k z^2 the delta-SPH term is a Laplacian:
    drho/dt ~= delta h c0 (lap rho)/2 = delta h c0 k (positive diffusion)."""
    k = 40.0
    c = _lattice_case(dim, dx=0.01, side=(0.2, 0.2, 0.2), rho_fn=lambda x: 1000.0 + k * (x[:, 2] - 0.4) ** 2)
    out = orc.eval_rates(c.params, c, mode=2)
    m = _interior(c, 2.5 * c.params.h)
    prm = c.params
    want = float(prm.delta) * float(prm.h) * float(prm.cs) * k
    np.testing.assert_allclose(out["drho"][m], want, rtol=0.05)






def _viscous_moment(prm, dim):
    """J = S_d / (d (d+2)) * int_0^2h r^(d+3) (-fr(r)) / (r^2 + eta^2) dr: the
    fourth moment of the artificial-viscosity weight, from its closed form
    (-fr = 5 adh (1 - q/2)^3 / h^2, adh = 21/(16 pi h^3) in 3-D and 7/(4 pi h^2)
    in 2-D); with eta = 0 it is exactly 1/(d+2)."""
    h = float(prm.h)
    eta = float(prm.eta)
    adh = 21.0 / (16.0 * np.pi * h ** 3) if dim == 3 else 7.0 / (4.0 * np.pi * h ** 2)
    r = np.linspace(0.0, 2.0 * h, 200_001)
    mfr = 5.0 * adh * (1.0 - r / (2.0 * h)) ** 3 / h ** 2
    f = r ** (dim + 3) * mfr / (r ** 2 + eta ** 2)
    integral = float(np.sum((f[1:] + f[:-1]) * np.diff(r)) / 2.0)
    surface = 4.0 * np.pi if dim == 3 else 2.0 * np.pi
    return surface * integral / (dim * (dim + 2))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("field,alpha,damper", [("shear", 0.5, 0.0), ("stretch", 0.5, 0.0), ("stretch", 0.0, 1.0)])
def test_viscosity_and_acoustic_damper_continuum_limits(orc, dim, field, alpha, damper):
    """Kernel B's fluid-fluid momentum terms (opt_src_v1/equation/ns.cu:210-213)
    at uniform rho0 and p = 0, against their continuum limits on a lattice
    with h = 3 dx, where the lattice's fourth moments are isotropic to a
    fraction of a percent (second-order Taylor expansion of the velocity):
      artificial viscosity  h c0 rho0 alpha / rho * J/2 * (2 grad(div v) + lap v)
      acoustic damper       h c0 rho0 damper / rho * grad(dofv),  dofv -> div v
    shear   v = (k z^2, 0, 0): div v = 0, lap v = (2k, 0, 0)    -> a_x = h c0 alpha J k
    stretch v = (k x^2, 0, 0): grad div v = lap v = (2k, 0, 0) -> a_x = h c0 (3 alpha J k + 2 damper k)
    A prefactor missing rho0, h or c0, a scaled alpha, a dropped factor 2 of
    grad(div v), or a flipped damper sign each miss by far more than 2 %."""
    k = 3.0
    if field == "shear":
        vf = lambda x: np.stack([k * (x[:, 2] - 0.5) ** 2, 0 * x[:, 0], 0 * x[:, 0]], 1)  # noqa: E731
    else:
        vf = lambda x: np.stack([k * (x[:, 0] - 1.2) ** 2, 0 * x[:, 0], 0 * x[:, 0]], 1)  # noqa: E731
    c = _lattice_case(dim, dx=0.01, side=(0.4, 0.4, 0.4), v_fn=vf, c0=20.0, h_over_dx=3.0)
    prm = c.params
    prm.afa = alpha
    prm.damper = damper
    out = orc.eval_rates(prm, c, mode=2)
    # the damper pairs dofv of neighbours, which need full support themselves
    m = _interior(c, 4.0 * float(prm.h) + 0.01)
    J = _viscous_moment(prm, dim)
    hc = float(prm.h) * float(prm.cs)  # rho0 / rho = 1
    want = hc * (alpha * J * k * (1.0 if field == "shear" else 3.0) + (2.0 * damper * k if field == "stretch" else 0))
    ax = out["acc"][m, 0]
    print(dim, field, alpha, damper, "J", J, "want", want, "got", ax.min(), ax.max())
    np.testing.assert_allclose(ax, want, rtol=0.02)
    # no cross-stream acceleration; gravity alone along z
    assert np.abs(out["acc"][m, 2] - float(prm.gravity)).max() < 0.02 * abs(want)
    if dim == 3:
        assert np.abs(out["acc"][m, 1]).max() < 0.02 * abs(want)


@pytest.mark.parametrize("dim", [2, 3])
def test_hydrostatic_pressure_balances_gravity(orc, dim):
    """-grad p / rho + g = 0 in the interior of a hydrostatic block."""
    H = 1.0
    c = _lattice_case(dim, dx=0.01, side=(0.2, 0.2, 0.2), p_fn=lambda x: 1000 * 9.81 * (H - x[:, 2]))
    out = orc.eval_rates(c.params, c, mode=2)
    m = _interior(c, 2.5 * c.params.h)
    assert np.abs(out["acc"][m, 2]).max() < 0.02 * 9.81
    assert np.abs(out["acc"][m, 0]).max() < 0.02 * 9.81


@pytest.mark.parametrize("dim", [2, 3])
def test_momentum_is_conserved(orc, dim):
    """Pressure, viscosity and damper forces act in equal and opposite pairs:
    sum_i m a_i = N m g for an isolated fluid blob with random state."""
    rng = np.random.default_rng(1)
    c = _lattice_case(dim, dx=0.02, side=(0.2, 0.2, 0.2))
    n = c.n
    c.pos = (c.pos + rng.uniform(-0.3, 0.3, (n, 3)) * 0.02).astype(np.float32)
    if dim == 2:
        c.pos[:, 1] = 0
    c.vel = rng.standard_normal((n, 3)).astype(np.float32)
    if dim == 2:
        c.vel[:, 1] = 0
    c.rho = rng.uniform(1000, 1010, n).astype(np.float32)
    c.p = rng.uniform(0, 2000, n).astype(np.float32)
    out = orc.eval_rates(c.params, c, mode=2)
    tot = out["acc"].astype(np.float64).sum(0) - np.array([0, 0, n * c.params.gravity])
    scale = np.abs(out["acc"]).sum(0) + 1e-30
    assert np.all(np.abs(tot) / scale < 1e-6), tot / scale


def test_galilean_invariance(orc):
    case = advanced(orc, case2d(0.02), 80, mode=2)
    # velocities on a 2^-16 grid so that the shifted values are exact in FP32
    case.vel = (np.round(case.vel.astype(np.float64) * 65536) / 65536).astype(np.float32)
    base = orc.eval_rates(case.params, case, mode=2)
    shifted = case.copy()
    fl = case.type == 1
    shifted.vel = case.vel.copy()
    shifted.vel[:, 0] += np.float32(0.375)  # walls move too: a uniform translation of the whole system
    shifted.vel[:, 2] -= np.float32(0.25)
    out = orc.eval_rates(case.params, shifted, mode=2)
    for k, mk in (("drho", "mag_drho"), ("acc", "mag_acc")):
        assert rel_err(out[k][fl], base[k][fl], base[mk][fl]) < 1e-5
    assert rel_err(out["dofv"], base["dofv"], base["mag_dofv"]) < 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_search_independent_of_grid_and_order(orc, dim):
    """A coarser cell grid and a shuffled particle order visit the same pairs:
    the results agree to rounding (a stencil that missed cells would not)."""
    case = advanced(orc, case2d(0.02) if dim == 2 else case3d(0.04), 60 if dim == 2 else 20, mode=2)
    base = orc.eval_rates(case.params, case, mode=2)
    coarse = case.copy()
    prm = coarse.params
    prm.gridSize = np.float32(prm.gridSize * 1.37)
    prm.gridxdim = int(np.ceil(prm.gridxdim / 1.37)) + 1
    prm.gridzdim = int(np.ceil(prm.gridzdim / 1.37)) + 1
    if dim == 3:
        prm.gridydim = int(np.ceil(prm.gridydim / 1.37)) + 1
    prm.numCells = prm.gridxdim * prm.gridydim * prm.gridzdim
    out = orc.eval_rates(prm, coarse, mode=2)
    # finer grids (edge h, h/2): the stencil reaches 2 and 4 cells on each side
    from tmp_sph_amd import sph
    fine = []
    for div in (2, 4):
        f = case.copy()
        f.params = sph.subdivide_params(case.params, div)
        assert f.params.numCells == case.params.numCells * div ** dim
        fine.append(orc.eval_rates(f.params, f, mode=2))
    rng = np.random.default_rng(2)
    perm = rng.permutation(case.n)
    sh = case.copy()
    sh.pos, sh.vel, sh.rho, sh.p, sh.type = case.pos[perm], case.vel[perm], case.rho[perm], case.p[perm], case.type[perm]
    o2 = orc.eval_rates(case.params, sh, mode=2)
    inv = np.argsort(perm)
    fl = case.type == 1
    for o, remap in ((out, np.arange(case.n)), (o2, inv), (fine[0], np.arange(case.n)), (fine[1], np.arange(case.n))):
        assert rel_err(o["drho"][remap][fl], base["drho"][fl], base["mag_drho"][fl]) < 1e-9
        assert rel_err(o["acc"][remap][fl], base["acc"][fl], base["mag_acc"][fl]) < 1e-9
        assert rel_err(o["p"][remap][~fl], base["p"][~fl], 1.0) < 1e-9


# ------------------------------------------------------------ integrator
def test_free_fall_is_exact_quadratic(orc):
    """An isolated particle has no neighbours: a = g, and the PC scheme
    advances x_{n+1} = x_n + v_n dt + a dt^2 / 2 exactly."""
    c = _lattice_case(2, dx=0.02, side=(0.02, 0.02, 0.02), origin=(1.0, 0.0, 0.9))
    assert c.n == 1
    prm = c.params
    steps = 120
    st, _, _, _ = orc.run(prm, c, steps, mode=2)
    t = steps * float(prm.timeStep)
    g = float(prm.gravity)
    z0 = float(c.pos[0, 2])
    assert st["pos64"][0, 2] == pytest.approx(z0 + 0.5 * g * t * t, rel=1e-12)
    assert st["vel"][0, 2] == pytest.approx(g * t, rel=1e-6)
    assert st["rho"][0] == pytest.approx(1000.0)


def test_escaped_particle_is_parked(orc):
    case = case2d(0.02)
    fl = np.flatnonzero(case.type == 1)
    case.pos[fl[3]] = (case.params.domain_max[0] + 0.003, 0.0, 0.2)
    case.vel[fl[3]] = (1.0, 0.0, 1.0)
    st, _, _, _ = orc.run(case.params, case, 1, mode=1)
    np.testing.assert_array_equal(st["pos"][fl[3]], np.asarray(case.params.park_pos, np.float32))
    assert np.all(st["vel"][fl[3]] == 0) and st["p"][fl[3]] == 0 and st["rho"][fl[3]] == 1000.0
    assert st["type"][fl[3]] == case.params.park_type


def test_tank_at_rest_stays_at_rest(orc):
    """Water filling the tank width is in hydrostatic balance: spurious
    velocities stay far below the gravity-wave speed."""
    c = case2d(0.02, tank=(1.0, 1.0, 1.0), column=(1.0, 1.0, 0.5))
    st, _, _, _ = orc.run(c.params, c, 300, mode=1)
    fl = c.type == 1
    assert np.abs(st["vel"][fl]).max() < 0.02 * np.sqrt(9.81 * 0.5)


def test_fp32_and_fp64_runs_agree(orc):
    c = case2d(0.02)
    r32, f32, h32, _ = orc.run(c.params, c, 300, mode=1, band=0.04, probe_every=50)
    r64, f64, h64, _ = orc.run(c.params, c, 300, mode=2, band=0.04, probe_every=50)
    np.testing.assert_allclose(f32, f64, atol=1e-5)
    np.testing.assert_allclose(h32, h64, atol=1e-5)
    fl = c.type == 1
    assert np.abs(r32["pos64"][fl] - r64["pos64"][fl]).max() < 1e-5


def test_dam_break_front_follows_the_classic_curve(orc):
    """Surge front of the 2-D dam break against the Martin & Moyce (1952)
    scaling: Z/a ~ 1.4 at T = t sqrt(2g/a) = 1 and ~ 1.9 at T = 1.5 (a is the
    column width).  Loose bounds: this pins the physics, not the digits."""
    c = case2d(0.02)
    prm = c.params
    a = 1.0
    T = np.array([1.0, 1.5])
    steps = (T / np.sqrt(2 * 9.81 / a) / prm.timeStep).astype(int)
    _, f1, _, _ = orc.run(prm, c, int(steps[1]), mode=1, band=0.04, probe_every=1)
    z = f1[steps - 1] / a
    assert 1.2 < z[0] < 1.6 and 1.6 < z[1] < 2.2, z
