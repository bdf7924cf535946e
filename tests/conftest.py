"""Shared fixtures.  GPU tests are marked ``gpu`` and run on an MI355X box;
everything else runs on the CPU (oracle, host logic, library loading)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def lib():
    from tmp_sph_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu(lib):
    """The HIP library on a visible device; fails (does not skip) without one."""
    from tmp_sph_amd import sph
    n = sph.device_count()
    assert n > 0, "no HIP device visible: GPU tests need the MI355X"
    return sph


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


def case2d(dx=0.02, **kw):
    from tmp_sph_amd import sph
    return sph.make_case(sph.dambreak_cfg(dim=2, dx=dx, **kw))


def case3d(dx=0.04, tank=(0.8, 0.4, 0.5), column=(0.3, 0.2, 0.25), **kw):
    from tmp_sph_amd import sph
    return sph.make_case(sph.dambreak_cfg(dim=3, dx=dx, tank=tank, column=column, **kw))


def advanced(orc, case, steps, mode=1):
    """The case advanced `steps` steps by the oracle (a non-trivial flow state)."""
    st, _, _, _ = orc.run(case.params, case, steps, mode=mode)
    out = case.copy()
    out.pos, out.vel, out.rho, out.p, out.type = st["pos"], st["vel"], st["rho"], st["p"], st["type"]
    return out


def rel_err(got, want, scale, slack=0.0):
    """max (|got - want| - slack)+ / max(|want|, scale), elementwise scale and
    slack allowed; slack is an absolute error budget (the oracle's sens_drho:
    what ulp-level input differences alone can cause)."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    if not got.size:
        return 0.0
    diff = np.maximum(np.abs(got - want) - np.asarray(slack, np.float64), 0.0)
    den = np.maximum(np.maximum(np.abs(want), scale), 1e-300)
    return float(np.max(np.where(diff == 0, 0.0, diff / den)))
