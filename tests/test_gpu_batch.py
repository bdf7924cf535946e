"""Batched mode: K uneven dam breaks stepped together (one launch per phase,
one hipGraph per step) must give every instance exactly the results of
running it alone; the autotuner's launch shapes must not change results."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from tmp_sph_amd import sph
    return [
        sph.make_case(sph.dambreak_cfg(dim=2, dx=0.02, column=(0.6, 1.0, 0.4))),
        sph.make_case(sph.dambreak_cfg(dim=2, dx=0.01, column=(0.3, 1.0, 0.5))),
        sph.make_case(sph.dambreak_cfg(dim=2, dx=0.025, column=(1.0, 1.0, 0.25))),
    ]


def test_batch_instances_bit_identical_to_single_runs(gpu):
    cases = _cases()
    steps = 12
    with gpu.Simulation(cases) as b:
        assert b.K == 3
        b.step(steps)
        sb = b.state()
        pb = b.pos64()
        fb, hb = b.probe_batch(0.04)
    for k, c in enumerate(cases):
        lo, hi = b.offsets[k], b.offsets[k + 1]
        with gpu.Simulation(c) as s:
            s.step(steps)
            ss = s.state()
            ps = s.pos64()
            f, h = s.probe(0.04)
        for key in ss:
            np.testing.assert_array_equal(sb[key][lo:hi], ss[key], err_msg=f"instance {k} {key}")
        np.testing.assert_array_equal(pb[lo:hi], ps)
        assert fb[k] == f and hb[k] == h
    # the instances advanced with their own time steps
    dts = {c.params.timeStep for c in cases}
    assert len(dts) == 3


def test_batch_graph_replay_matches_eager(gpu):
    cases = _cases()
    with gpu.Simulation(cases) as a, gpu.Simulation(cases) as b:
        a.step(5)
        b.capture()
        b.replay(5)
        b.synchronize()
        np.testing.assert_array_equal(a.pos64(), b.pos64())
        sa, sb = a.state(), b.state()
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k])


@pytest.mark.skip(reason="faulted the GPU twice in one session: the autotuner ran the neighbour-list kernels "
                         "with a copy of the work arrays taken before the lists were allocated (null list "
                         "pointer); fixed (lists are no longer autotune candidates, list launches check the "
                         "copy they get) but, as the retest of a fault fix faulted, kept off the GPU until "
                         "the next session")
def test_autotune_keeps_results(gpu):
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    case = make_case(dambreak_cfg(dim=3, dx=0.04, tank=(0.8, 0.4, 0.5), column=(0.3, 0.2, 0.25)))
    with gpu.Simulation(case) as a, gpu.Simulation(case) as b:
        best, us = b.autotune(reps=2)
        assert best["nbr_block"] in (64, 128, 256) and best["nbr_unroll"] in (1, 2, 4, 8) and us > 0
        a.step(4)
        b.step(4)
        np.testing.assert_array_equal(a.pos64(), b.pos64())
        for blk, un, rm in ((64, 1, 0), (128, 8, 1), (192, 2, 0)):
            with gpu.Simulation(case) as c:
                c.set_tuning(nbr_block=blk, nbr_unroll=un, xcd_remap=rm)
                c.step(4)
                np.testing.assert_array_equal(a.pos64(), c.pos64())


def test_autotune_refuses_while_disabled(gpu):
    """The launch-shape search is compiled out (it faulted the GPU with the
    cause unresolved): the entry point reports a state error before any
    launch and leaves the context's shape as it was."""
    from tmp_sph_amd import _lib
    case = gpu.make_case(gpu.dambreak_cfg(dim=2, dx=0.02))
    with gpu.Simulation(case) as s:
        before = s.tuning.as_dict()
        with pytest.raises(_lib.SphError, match="disabled"):
            s.autotune(reps=1)
        assert s.tuning.as_dict() == before


@pytest.mark.parametrize("lds", [3, 4])
def test_xcd_remap_only_moves_tiles_between_blocks(gpu, lds):
    """Every tile-to-block map (identity, one run per XCD, runs of G tiles
    dealt to the XCDs in turn, with a partial last round) computes the same
    tiles: bit-identical steps, single and batched."""
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    cases = [make_case(dambreak_cfg(dim=3, dx=0.03, tank=(0.8, 0.4, 0.5), column=(0.3, 0.4, 0.3))), _cases()]
    for c in cases:
        ref = None
        for rm in (0, 1, 2, 16, 64, 4096):
            with gpu.Simulation(c) as s:
                s.set_tuning(nbr_lds=lds, xcd_remap=rm)
                s.step(4)
                r = s.pos64()
            if ref is None:
                ref = r
                continue
            np.testing.assert_array_equal(r, ref, err_msg=f"nbr_lds={lds} xcd_remap={rm}")
    with gpu.Simulation(cases[0]) as s:
        with pytest.raises(Exception, match="power of two"):
            s.set_tuning(xcd_remap=48)


_VARIANTS = ([(lds, compact, pf, 0, un) for lds in (0, 1) for compact in (0, 1) for pf in (0, 1) for un in (1, 4, 8)]
             + [(2, 0, pf, 0, un) for pf in (0, 1) for un in (1, 4, 8)]
             + [(lds, 2, 0, 0, un) for lds in (0, 2, 3) for un in (1, 2, 8)]
             + [(3, 0, pf, 0, un) for pf in (0, 1) for un in (1, 4, 8)]
             + [(0, 0, 0, 1, un) for un in (1, 4, 8)]
             + [(4, 0, pf, 0, un) for pf in (0, 1) for un in (1, 2, 4)])


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_loop_variants_bit_identical(gpu, dim):
    """LDS staging (whole stencil, plane by plane, row by row with the
    neighbour data), the scan/dense (compact)
    loop, mask groups, prefetched neighbour loads and
    neighbour lists visit the same pairs in the same order as the plain global
    loop: every rate is bit-identical."""
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    if dim == 3:
        case = make_case(dambreak_cfg(dim=3, dx=0.03, tank=(0.8, 0.4, 0.5), column=(0.3, 0.4, 0.3)))
    else:
        case = make_case(dambreak_cfg(dim=2, dx=0.01))
    ref = None
    for lds, compact, pf, lst, un in _VARIANTS:
        with gpu.Simulation(case) as s:
            s.step(3)  # default loop; leaves the lattice so every rate is non-trivial
            s.set_tuning(nbr_lds=lds, nbr_compact=compact, nbr_prefetch=pf, nbr_list=lst, nbr_unroll=un)
            r = s.eval_rates()
            s.step(2)
            r["pos"] = s.pos64()
        if ref is None:
            ref = r
            continue
        for k in ("drho", "acc", "dofv", "p", "rho", "pos"):
            np.testing.assert_array_equal(r[k], ref[k],
                                          err_msg=f"lds={lds} compact={compact} prefetch={pf} list={lst} "
                                                  f"unroll={un} {k}")


def test_neighbour_loop_variants_batched(gpu):
    """Waves that straddle instances take the per-instance path of every
    variant; results stay identical to the plain loop."""
    cases = _cases()
    ref = None
    for lds, compact, pf, lst, un in _VARIANTS:
        with gpu.Simulation(cases) as b:
            b.set_tuning(nbr_lds=lds, nbr_compact=compact, nbr_prefetch=pf, nbr_list=lst, nbr_unroll=un)
            b.step(6)
            r = b.pos64()
        if ref is None:
            ref = r
            continue
        np.testing.assert_array_equal(r, ref, err_msg=f"lds={lds} compact={compact} prefetch={pf} list={lst} "
                                                     f"unroll={un}")


@pytest.mark.parametrize("dim", [2, 3])
def test_fine_rows_on_the_halved_grid_bit_identical_across_shapes(gpu, dim):
    """On the halved grid every fine-row launch shape (unroll, prefetch,
    XCD remap) gives bit-identical rates and steps; a batch gives every
    instance the results of running it alone."""
    from tmp_sph_amd.sph import dambreak_cfg, make_case
    if dim == 3:
        cases = [make_case(dambreak_cfg(dim=3, dx=0.03, tank=(0.8, 0.4, 0.5), column=(0.3, 0.4, 0.3)))]
    else:
        cases = _cases()
    ref = None
    for un, pf, rm in ((1, 0, 0), (2, 1, 0), (4, 0, 1), (4, 1, 1)):
        with gpu.Simulation(cases) as s:
            s.set_tuning(nbr_lds=4, cell_div=2, nbr_unroll=un, nbr_prefetch=pf, xcd_remap=rm)
            s.step(3)
            r = s.eval_rates()
            s.step(2)
            r["pos"] = s.pos64()
        if ref is None:
            ref = r
            continue
        for k in ("drho", "acc", "dofv", "p", "rho", "pos", "keys", "perm"):
            np.testing.assert_array_equal(r[k], ref[k], err_msg=f"unroll={un} prefetch={pf} remap={rm} {k}")
    if len(cases) > 1:
        with gpu.Simulation(cases) as b:
            b.set_tuning(nbr_lds=4, cell_div=2)
            b.step(6)
            pb = b.pos64()
            offs = b.offsets
        for k, c in enumerate(cases):
            with gpu.Simulation(c) as s:
                s.set_tuning(nbr_lds=4, cell_div=2)
                s.step(6)
                np.testing.assert_array_equal(pb[offs[k]:offs[k + 1]], s.pos64(), err_msg=f"instance {k}")


def test_c4_sweep_keys_order_and_members(gpu, orc):
    """The 64-member batched sweep (uneven 2-D dam breaks, seeded): keys and
    sorted order of the whole concatenated batch bit-exact against the
    oracle's per-instance hash plus instance cell offsets under a stable
    sort; after five replayed steps of the captured batch step, the smallest
    and the largest member and one of spacing 0.002 m are bit-identical to
    the same cases run alone."""
    from tmp_sph_amd import cases as cs
    sweep = cs.sweep_cases()
    members = cs.sweep_members()
    assert len(sweep) == 64
    steps = 5
    with gpu.Simulation(sweep) as b:
        r = b.eval_rates()
        b.capture()
        b.replay(steps)
        b.synchronize()
        sb = b.state()
        pb = b.pos64()
        offsets = b.offsets
    keys, off = [], 0
    for c in sweep:
        h, _ = orc.calcHash(c.params, c.pos)
        keys.append(h.astype(np.int64) + off)
        off += c.params.numCells
    keys = np.concatenate(keys)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(r["keys"], keys[order])
    np.testing.assert_array_equal(r["perm"], order)
    sizes = [c.n for c in sweep]
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)),
                   next(k for k, m in enumerate(members) if m.dx == 0.002)})
    for k in pick:
        lo, hi = offsets[k], offsets[k + 1]
        with gpu.Simulation(sweep[k]) as s:
            s.step(steps)
            ss = s.state()
            ps = s.pos64()
        for key in ss:
            np.testing.assert_array_equal(sb[key][lo:hi], ss[key], err_msg=f"member {k} {key}")
        np.testing.assert_array_equal(pb[lo:hi], ps, err_msg=f"member {k} pos64")
