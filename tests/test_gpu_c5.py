"""The learned surrogate's evaluation geometry at full size: a 0.8 x 0.5 m
column in the 4.0 x 1.0 m tank at dx = 0.002 m (109 018 particles, 100 000
fluid), held out of the training sweep.  The radius graph is the SPH pair set
(against an independent KD-tree search), and a rollout whose judge rejects
every network step is bit-identical to the plain run."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _c5():
    from tmp_sph_amd import sph
    return sph.make_case(sph.dambreak_cfg(dim=2, dx=0.002, tank=(4.0, 1.0, 1.0), column=(0.8, 1.0, 0.5)))


def test_c5_radius_graph_is_the_pair_set(gpu):
    from scipy.spatial import cKDTree
    from tmp_sph_amd.surrogate import Gnn
    case = _c5()
    n = case.n
    assert n == 109018 and case.n_fluid == 100000
    g = Gnn(layers=2, max_nodes=n, max_edges=30 * n, train=False)
    g.load(case.params, case.pos, np.zeros((5, n, 3), np.float32), case.rho, case.type)
    off, send, recv = g.graph()
    g.close()
    # candidates from a KD tree over a slightly larger radius, kept by the
    # FP32 squared distance (unfused sum) the solver's pair test uses
    p = case.pos[:, [0, 2]].astype(np.float32)
    cut = np.float32(case.params.rr_cut)
    pairs = cKDTree(p.astype(np.float64)).query_pairs(float(np.sqrt(cut)) * 1.0001, output_type="ndarray")
    d = p[pairs[:, 0]] - p[pairs[:, 1]]
    pairs = pairs[d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] <= cut]
    assert off[0] == 0 and off[-1] == send.size == recv.size == 2 * len(pairs)
    # CSR by receiver
    assert np.all(np.diff(off) >= 0)
    np.testing.assert_array_equal(recv, np.repeat(np.arange(n), np.diff(off)))
    got = np.sort(recv.astype(np.int64) * n + send)
    exp = np.sort(np.concatenate([pairs[:, 0].astype(np.int64) * n + pairs[:, 1],
                                  pairs[:, 1].astype(np.int64) * n + pairs[:, 0]]))
    np.testing.assert_array_equal(got, exp)


def test_c5_rejected_intervals_reproduce_the_plain_run(gpu):
    """Seven intervals of 10 steps: the first four fill the velocity history
    (with the start state, five frames), the network steps of the other three
    are rejected (dev_tol < 0) and redone with plain steps; positions (FP64),
    state and probes equal 70 plain steps bit for bit.  The judge's density
    readings of those plain intervals stay inside the rollout's 1 % bound."""
    from tmp_sph_amd import sph
    from tmp_sph_amd.surrogate import Gnn
    case = _c5()
    k, intervals = 10, 7
    g = Gnn(layers=10, max_nodes=case.n, max_edges=30 * case.n, train=False, seed=3)
    with sph.Simulation(case) as a, sph.Simulation(case) as b:
        st, f, h = g.rollout(a, k, intervals, dev_tol=-1.0)
        print("rollout:", st)
        assert st["accepted"] == 0 and st["fallbacks"] == intervals - 4
        assert st["plain_steps"] == k * intervals
        b.step(k * intervals)
        np.testing.assert_array_equal(a.pos64(), b.pos64())
        sa, sb = a.state(), b.state()
        for key in sa:
            np.testing.assert_array_equal(sa[key], sb[key], err_msg=key)
        fb, hb = b.probe(0.02)
        assert f[-1] == fb and h[-1] == hb
    assert st["plain_max_dev"] <= 0.01, st
    g.close()
