"""sph_sortParticles_noalloc -- the reference-surface sort without any device
allocation (scratch in a file-scope device array) -- against
sph_sortParticles, which allocates its scratch on first use: bit-identical
keys and indices on two calls in a row, the second smaller, and both equal
to numpy's stable argsort."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sorted_pair(call_name, keys):
    from tmp_sph_amd import _lib, sph
    n = keys.size
    dk = sph.DeviceArray.from_host(keys.astype(np.int32))
    di = sph.DeviceArray.from_host(np.arange(n, dtype=np.int32))
    _lib.call(call_name, dk.c, di.c, int(n))
    sph.synchronize()
    return dk.numpy(), di.numpy()


def test_noalloc_sort_matches_the_allocating_sort(gpu):
    from tmp_sph_amd import _lib, sph
    case = sph.make_case(sph.dambreak_cfg(dim=3, dx=0.01, tank=(1.6, 0.6, 0.8), column=(0.6, 0.6, 0.4)))
    prm = case.params
    sph.set_params(prm)
    rng = np.random.default_rng(7)
    # a real case's cell keys in a shuffled order, then a smaller random set with many ties
    pos = sph.DeviceArray.from_host(case.pos.reshape(-1))
    h = sph.DeviceArray((case.n,), np.int32)
    i = sph.DeviceArray((case.n,), np.int32)
    sph.calcHash(pos, h, i, case.n)
    keys1 = h.numpy()[rng.permutation(case.n)]
    keys2 = rng.integers(0, 97, 12_345).astype(np.int32)
    assert keys2.size < keys1.size
    for keys in (keys1, keys2):
        k_a, i_a = _sorted_pair("sph_sortParticles", keys)
        k_n, i_n = _sorted_pair("sph_sortParticles_noalloc", keys)
        np.testing.assert_array_equal(k_n, k_a)
        np.testing.assert_array_equal(i_n, i_a)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(i_n, order)
        np.testing.assert_array_equal(k_n, keys[order])


def test_noalloc_sort_refuses_more_than_its_scratch(gpu):
    from tmp_sph_amd import _lib, sph
    sph.set_params(sph.reference_params())
    d = sph.DeviceArray((4,), np.int32)
    limit = 1 << 22  # SPH_NOALLOC_MAX_PARTICLES
    with pytest.raises(_lib.SphError) as e:
        _lib.call("sph_sortParticles_noalloc", d.c, d.c, limit + 1)
    assert e.value.code == _lib.SPH_E_INVAL
