"""GPU checks of the file paths around the step: checkpoint / restart, the
asynchronous frame writer, and a case loaded from the reference's .dat
initial-state files."""
from __future__ import annotations


import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(dim, dx=0.02, **kw):
    from tmp_sph_amd import sph
    if dim == 2:
        return sph.make_case(sph.dambreak_cfg(dim=2, dx=dx, **kw))
    return sph.make_case(sph.dambreak_cfg(dim=3, dx=0.04, tank=(1.6, 0.5, 0.6), column=(0.5, 0.5, 0.3)))


@pytest.mark.parametrize("dim", [2, 3])
def test_restart_continues_bit_identically(gpu, tmp_path, dim):
    from tmp_sph_amd import sph
    case = _case(dim)
    ck = tmp_path / "state.ckpt"
    with sph.Simulation(case) as a:
        a.step(37)
        a.checkpoint(ck)
        a.step(41)
        want = a.state()
        want64 = a.pos64()
    with sph.Simulation.restore(ck) as b:
        assert b.n == case.n and b.n_fluid == case.n_fluid
        assert b.steps_done == 37
        b.step(41)
        got = b.state()
        np.testing.assert_array_equal(b.pos64(), want64)
        assert b.steps_done == 78
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_restart_of_a_batch_and_under_graph_replay(gpu, tmp_path):
    from tmp_sph_amd import sph
    cases = [_case(2, dx=0.02), _case(2, dx=0.025, column=(0.4, 1.0, 0.4))]
    ck = tmp_path / "batch.ckpt"
    with sph.Simulation(cases) as a:
        a.step(10)
        a.checkpoint(ck)
        a.capture()
        a.replay(20)
        want = a.state()
    with sph.Simulation.restore(ck) as b:
        assert b.K == 2 and list(b.offsets) == [0, cases[0].n, cases[0].n + cases[1].n]
        b.step(20)
        got = b.state()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_restore_rejects_a_damaged_checkpoint(gpu, tmp_path):
    from tmp_sph_amd import sph
    from tmp_sph_amd._lib import SphError
    ck = tmp_path / "s.ckpt"
    with sph.Simulation(_case(2)) as a:
        a.step(3)
        a.checkpoint(ck)
    raw = bytearray(ck.read_bytes())
    bad = tmp_path / "bad.ckpt"
    raw[len(raw) // 2] ^= 0x10
    bad.write_bytes(bytes(raw))
    with pytest.raises(SphError, match="checksum mismatch"):
        sph.Simulation.restore(bad)
    bad.write_bytes(ck.read_bytes()[:-9])
    with pytest.raises(SphError, match="truncated"):
        sph.Simulation.restore(bad)


@pytest.mark.parametrize("binary", [True, False])
def test_async_frames_equal_synchronous_writes(gpu, tmp_path, binary):
    from tmp_sph_amd import files, sph
    case = _case(2)
    out, ref = tmp_path / "async", tmp_path / "sync"
    ref.mkdir()
    with sph.Simulation(case) as s:
        s.output_open(out, binary=binary, rho=True, type=True)
        for frame in range(6):
            s.step(7)
            s.output_frame(frame)  # returns before the frame is on disk
            st = s.state()
            files.write_vtk(ref / f"fluid_t={frame}.vtk", st["pos"], st["vel"], st["p"], rho=st["rho"],
                            type=st["type"], binary=binary)
        assert s.output_flush() == 6
        s.output_close()
    for frame in range(6):
        name = f"fluid_t={frame}.vtk"
        assert (out / name).read_bytes() == (ref / name).read_bytes(), name


def test_async_frames_overlap_steps_without_changing_them(gpu, tmp_path):
    """Frames queued back to back (both buffers busy) while steps run hold the
    states of their own steps; the steps equal a run without output."""
    from tmp_sph_amd import files, sph
    case = _case(2)
    with sph.Simulation(case) as a, sph.Simulation(case) as b:
        a.output_open(tmp_path, binary=True)
        for frame in range(5):
            a.step(3)
            a.output_frame(frame)
        a.step(3)
        a.output_close()
        b.step(12)
        mid = b.state()
        b.step(6)
        np.testing.assert_array_equal(a.state()["pos"], b.state()["pos"])
    d = files.read_vtk(tmp_path / "fluid_t=3.vtk")
    np.testing.assert_array_equal(d["points"], mid["pos"])
    np.testing.assert_array_equal(d["p"], mid["p"])


def test_async_write_errors_are_reported(gpu, tmp_path):
    from tmp_sph_amd import _lib, sph
    from tmp_sph_amd._lib import SphError
    with sph.Simulation(_case(2)) as s:
        # a directory that does not exist: the writer thread's fopen fails
        _lib.call("sph_sim_output_open", s._h, str(tmp_path / "missing").encode(), 1, 0)
        s.output_frame(0)
        with pytest.raises(SphError, match="cannot create"):
            s.output_flush()
        with pytest.raises(SphError, match="cannot create"):
            s.output_frame(1)
        with pytest.raises(SphError):
            s.output_close()
        with pytest.raises(SphError, match="not open"):
            s.output_frame(2)
        # the context itself still steps
        s.step(2)
    # destroying a context with frames still queued drains them first
    with sph.Simulation(_case(2)) as s:
        s.output_open(tmp_path / "d", binary=False)
        s.output_frame(0)
        s.output_frame(1)
    assert (tmp_path / "d" / "fluid_t=1.vtk").stat().st_size > 0


def test_case_loaded_from_dat_files_runs_like_the_generated_one(gpu, tmp_path):
    from tmp_sph_amd import files, sph
    case = _case(2)
    files.write_dat(tmp_path, case.pos, case.vel, case.rho, case.p, case.type)
    d = files.read_dat(tmp_path)
    loaded = sph.Case(case.params.copy(), d["pos"], d["vel"], d["rho"], d["p"], d["type"],
                      int(np.count_nonzero(d["type"] == 1)))
    with sph.Simulation(case) as a, sph.Simulation(loaded) as b:
        a.step(20)
        b.step(20)
        np.testing.assert_array_equal(a.pos64(), b.pos64())


def test_contexts_leave_the_reference_surface_parameters_alone(gpu, tmp_path):
    """Creating or restoring a context validates its instances' parameters
    without replacing the process-wide block of the reference call surface
    (sph_set_params / sph_get_params)."""
    from tmp_sph_amd import sph
    a = sph.reference_params()
    sph.set_params(a)
    case = _case(2)
    assert bytes(case.params) != bytes(a)
    ck = tmp_path / "p.ckpt"
    with sph.Simulation(case) as s:
        s.step(2)
        s.checkpoint(ck)
    assert bytes(sph.get_params()) == bytes(a)
    with sph.Simulation.restore(ck):
        pass
    assert bytes(sph.get_params()) == bytes(a)
