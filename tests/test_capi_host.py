"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/sph_capi.h declares, and its host-only entry points (parameter
builders, case generator, argument validation) behave."""
from __future__ import annotations

import ctypes as C
import math
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "sph_capi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared in sph_capi.h but not exported"


def test_binding_covers_the_header():
    from tmp_sph_amd import _lib
    assert set(_declared()) == set(_lib.EXPORTS)


def test_abi_sizes(lib):
    from tmp_sph_amd._lib import DamBreakCfg, SphParams, SphTuning
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.sph_abi_sizes(C.byref(a), C.byref(b), C.byref(c)) == 0
    assert (a.value, b.value, c.value) == (C.sizeof(SphParams), C.sizeof(DamBreakCfg), C.sizeof(SphTuning))


def test_reference_params_recipe(lib):
    """Derived values of src/sph/sph_run.cu:60-101 (dx = 0.01)."""
    from tmp_sph_amd import sph
    p = sph.reference_params()
    f32 = np.float32
    assert p.h == f32(0.015) and p.kh == f32(0.03) and p.gridSize == p.kh
    assert p.eta == pytest.approx(0.0015, rel=1e-6)
    assert p.delta == f32(0.2) and p.afa == f32(0.05) and p.cs == 25.0 and p.timeStep == f32(2e-4)
    assert p.particleMass == pytest.approx(1e-3, rel=1e-6)
    assert p.adh == pytest.approx(21.0 / (16.0 * 3.141592654 * 0.015 ** 3), rel=1e-6)
    # FP32 (max - min) / size: 3.421/0.03 -> 115; 1.2/0.03 lands just above 40 in FP32 -> 41
    assert (p.gridxdim, p.gridydim, p.gridzdim) == (115, 41, 41)
    assert p.numCells == 115 * 41 * 41
    assert p.hash_max == 115 * 41 * 41 + 115 * 41 + 115
    assert p.eos == 0 and list(p.park_pos) == [0.0, 0.5, pytest.approx(-0.08)]
    assert list(p.domain_max) == [3.23, 1.0, 1.0]


@pytest.mark.parametrize("dim,dx,tank,column,nf,nw", [
    (2, 0.01, (4.0, 1.0, 1.0), (1.0, 1.0, 0.5), 5000, (400 + 6) * 3 + 2 * 100 * 3),
    (2, 0.001, (4.0, 1.0, 1.0), (1.0, 1.0, 0.5), 500_000, (4000 + 6) * 3 + 2 * 1000 * 3),
    (3, 0.005, (3.2, 1.0, 1.0), (1.0, 0.5, 0.5), 2_000_000,
     646 * 206 * 3 + 2 * 646 * 200 * 3 + 2 * 200 * 200 * 3),
])
def test_case_counts(lib, dim, dx, tank, column, nf, nw):
    from tmp_sph_amd import sph
    nt, f = sph.case_counts(sph.dambreak_cfg(dim=dim, dx=dx, tank=tank, column=column))
    assert f == nf and nt - f == nw


def test_dambreak_params(lib):
    from tmp_sph_amd import sph
    p = sph.make_case(sph.dambreak_cfg(dim=2, dx=0.01)).params
    c0 = 10 * math.sqrt(2 * 9.81 * 0.5)
    assert p.cs == pytest.approx(c0, rel=1e-6)
    assert p.timeStep == pytest.approx(0.2 * 0.013 / c0, rel=1e-6)
    assert p.h == pytest.approx(0.013, rel=1e-7) and p.gridSize == pytest.approx(0.026, rel=1e-7)
    assert p.particleMass == pytest.approx(0.1, rel=1e-6)  # dx^2 rho0
    assert p.tait_B == pytest.approx(1000 * c0 * c0 / 7, rel=1e-6)
    assert p.gridydim == 1 and p.gridymin == 0.0
    p3 = sph.make_case(sph.dambreak_cfg(dim=3, dx=0.04, tank=(0.8, 0.4, 0.5), column=(0.3, 0.2, 0.25))).params
    assert p3.particleMass == pytest.approx(0.04 ** 3 * 1000, rel=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_wendland_normalisation(lib, dim):
    """The kernel constant makes the integral of W over its support 1."""
    from tmp_sph_amd import sph
    cfg = sph.dambreak_cfg(dim=dim, dx=0.01, tank=(1, 1, 1), column=(0.5, 0.5, 0.5))
    p = sph.make_case(cfg).params if dim == 2 else None
    if p is None:
        p = C_params(cfg)
    h = float(p.h)
    q = np.linspace(0.0, 2.0, 200_001)
    w = float(p.adh) * (1 - q / 2) ** 4 * (2 * q + 1)
    r = q * h
    integrand = w * (2 * np.pi * r if dim == 2 else 4 * np.pi * r * r)
    total = np.trapz(integrand, r)
    assert total == pytest.approx(1.0, rel=2e-6)


def C_params(cfg):
    from tmp_sph_amd._lib import SphParams, call
    p = SphParams()
    call("sph_params_dambreak", C.byref(cfg), C.byref(p))
    return p


def test_cutoff_threshold_is_the_reference_test(lib):
    """rr <= rr_cut selects exactly the pairs with FP32 sqrtf(rr)/h <= 2."""
    from tmp_sph_amd import sph
    for dx in (0.01, 0.001, 0.005, 0.0137):
        p = sph.make_case(sph.dambreak_cfg(dim=2, dx=dx)).params if dx != 0.005 else sph.reference_params(0.005)
        h = np.float32(p.h)
        rc = np.float32(p.rr_cut)
        assert np.sqrt(rc) / h <= np.float32(2.0)
        up = np.nextafter(rc, np.float32(np.inf))
        assert np.sqrt(up) / h > np.float32(2.0)


def test_case_geometry(lib):
    from tmp_sph_amd import sph
    case = sph.make_case(sph.dambreak_cfg(dim=2, dx=0.02))
    p = case.params
    fl = case.type == 1
    x, z = case.pos[:, 0], case.pos[:, 2]
    assert np.all(case.pos[:, 1] == 0)
    assert x[fl].min() == pytest.approx(0.01) and x[fl].max() == pytest.approx(0.99)
    assert z[fl].max() == pytest.approx(0.49)
    # walls: outside the tank faces, three layers
    wx, wz = x[~fl], z[~fl]
    assert np.all((wz < 0) | (wx < 0) | (wx > 4.0))
    assert set(np.round(wz[wz < 0] / 0.02, 3)) == {-0.5, -1.5, -2.5}
    # no two particles share a lattice site
    keys = np.round(case.pos / 0.01).astype(np.int64)
    assert len(np.unique(keys, axis=0)) == case.n
    # hydrostatic pressure and Tait density
    H = 0.5
    np.testing.assert_allclose(case.p[fl], 1000 * 9.81 * (H - z[fl]), rtol=1e-6)
    B = float(p.tait_B)
    np.testing.assert_allclose(case.rho[fl], 1000 * (1 + case.p[fl].astype(np.float64) / B) ** (1 / 7), rtol=1e-6)
    assert np.all(case.p[~fl] == 0) and np.all(case.rho[~fl] == 1000)


def test_argument_validation(lib):
    from tmp_sph_amd import _lib, sph
    bad = sph.reference_params()
    bad.dim = 4
    with pytest.raises(_lib.SphError) as e:
        sph.set_params(bad)
    assert e.value.code == _lib.SPH_E_INVAL
    bad = sph.reference_params()
    bad.numCells += 1
    with pytest.raises(_lib.SphError):
        sph.set_params(bad)
    bad = sph.reference_params()
    bad.gridSize = bad.h  # smaller than 2h: the 3-cell stencil would miss pairs
    with pytest.raises(_lib.SphError):
        sph.set_params(bad)
    with pytest.raises(_lib.SphError):
        sph.make_case(sph.dambreak_cfg(dim=2, dx=0.01, column=(5.0, 1.0, 0.5)))  # column wider than tank
    h = C.c_void_p()
    assert lib.sph_sim_create(None, 10, None, None, None, None, None, C.byref(h)) == _lib.SPH_E_INVAL
