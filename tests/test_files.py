"""Host-side file formats (no GPU): the initial-state .dat directory of the
reference's ini_fluid (src/io/io.cu:6-111) and its legacy-VTK frames (ASCII
src/io/io.cu:164-196, big-endian binary opt_src_v3/io/io.cu:121-200)."""
from __future__ import annotations

import numpy as np
import pytest

from tmp_sph_amd import files
from tmp_sph_amd._lib import SphError


def _arrays(n, seed=0):
    rng = np.random.default_rng(seed)
    pos = (rng.standard_normal((n, 3)) * [1.0, 1e-3, 30.0]).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    vel[0] = [0.0, -0.0, 1e-30]
    vel[1] = [3.4028235e38, -1.1754944e-38, 1.4e-45]  # largest, smallest normal, smallest subnormal
    rho = (1000 + rng.standard_normal(n)).astype(np.float32)
    p = (rng.standard_normal(n) * 1e4).astype(np.float32)
    typ = rng.integers(0, 3, n).astype(np.int32)
    zone = rng.integers(-5, 5, n).astype(np.int32)
    acc = rng.standard_normal((n, 3)).astype(np.float32)
    return pos, vel, acc, rho, p, typ, zone


def test_dat_round_trip_is_exact(tmp_path, lib):
    pos, vel, acc, rho, p, typ, zone = _arrays(257)
    files.write_dat(tmp_path, pos, vel, rho, p, typ, acc=acc, zone=zone)
    assert files.count_dat(tmp_path) == 257
    got = files.read_dat(tmp_path)
    for k, want in dict(pos=pos, vel=vel, acc=acc, rho=rho, p=p, type=typ, zone=zone).items():
        np.testing.assert_array_equal(got[k], want, err_msg=k)
        assert got[k].dtype == want.dtype
    # the optional fields are left out when not asked for
    part = files.read_dat(tmp_path, acc=False, zone=False)
    assert "acc" not in part and "zone" not in part


def test_dat_reads_the_reference_text_layout(tmp_path, lib):
    """Files as a generator for the reference writes them: one particle per
    line, any whitespace, integers or scientific notation for floats."""
    (tmp_path / "h_pos.dat").write_text("0 0.005 1e-2\n  2.5E+0\t-0.125 3\n")
    (tmp_path / "h_vel.dat").write_text("0 0 0\n0 0 -1.5\n")
    (tmp_path / "h_ac.dat").write_text("0 0 -9.81\n0 0 -9.81\n")
    (tmp_path / "rhop.dat").write_text("1000\n1000.5\n")
    (tmp_path / "p.dat").write_text("0\n12.75\n")
    (tmp_path / "particle_type.dat").write_text("0\n1\n")
    (tmp_path / "particle_zone.dat").write_text("0\n2\n")
    d = files.read_dat(tmp_path)
    # every token here is exactly representable in float32
    np.testing.assert_array_equal(d["pos"], np.array([[0, 0.005, 0.01], [2.5, -0.125, 3]], np.float32))
    np.testing.assert_array_equal(d["vel"][1], [0, 0, -1.5])
    np.testing.assert_array_equal(d["acc"][:, 2], np.float32(-9.81))
    np.testing.assert_array_equal(d["rho"], [1000, 1000.5])
    np.testing.assert_array_equal(d["p"], [0, 12.75])
    np.testing.assert_array_equal(d["type"], [0, 1])
    np.testing.assert_array_equal(d["zone"], [0, 2])


def test_dat_errors_name_the_file(tmp_path, lib):
    pos, vel, acc, rho, p, typ, zone = _arrays(8)
    files.write_dat(tmp_path, pos, vel, rho, p, typ)
    (tmp_path / "p.dat").write_text("1 2 3\n")  # short
    with pytest.raises(SphError, match=r"p\.dat: expected 8 values, read 3"):
        files.read_dat(tmp_path)
    (tmp_path / "h_vel.dat").unlink()
    with pytest.raises(SphError, match=r"cannot open .*h_vel\.dat"):
        files.read_dat(tmp_path)
    with pytest.raises(SphError, match=r"cannot open .*rhop\.dat"):
        files.count_dat(tmp_path / "missing")
    files.write_dat(tmp_path, pos[:4], vel[:4], rho[:4], p[:4], typ[:4])
    (tmp_path / "rhop.dat").write_text("1 2 x 4\n")  # not a number: stops there
    with pytest.raises(SphError, match=r"rhop\.dat: expected 4 values, read 2"):
        files.read_dat(tmp_path, 4)


def _reference_ascii(pos, vel, p) -> bytes:
    """The byte stream of the reference's ASCII writer for these arrays: a
    stream in scientific mode with precision 6 prints as "%.6e"."""
    n = len(p)
    out = ["# vtk DataFile Version 3.0", "sph data", "ASCII", "DATASET UNSTRUCTURED_GRID", f"POINTS {n} double"]
    out += ["%.6e %.6e %.6e" % tuple(float(x) for x in r) for r in pos]
    out += [f"POINT_DATA {n}", "SCALARS p double 1", "LOOKUP_TABLE DEFAULT"]
    out += ["%.6e" % float(x) for x in p]
    out += ["VECTORS velocity double"]
    out += ["%.6e %.6e %.6e" % tuple(float(x) for x in r) for r in vel]
    return ("\n".join(out) + "\n").encode()


def _reference_binary(pos, vel, p) -> bytes:
    """The reference's binary writer: big-endian float32 words, each block
    followed directly by the next keyword line."""
    n = len(p)
    b = f"# vtk DataFile Version 3.0\nsph data\nBINARY\nDATASET UNSTRUCTURED_GRID\nPOINTS {n} float\n".encode()
    b += np.asarray(pos, ">f4").tobytes()
    b += f"POINT_DATA {n}\nSCALARS p float 1\nLOOKUP_TABLE DEFAULT\n".encode()
    b += np.asarray(p, ">f4").tobytes()
    b += b"VECTORS velocity float\n"
    b += np.asarray(vel, ">f4").tobytes()
    return b


@pytest.mark.parametrize("binary", [False, True])
def test_vtk_bytes_match_the_reference_writer(tmp_path, lib, binary):
    pos, vel, acc, rho, p, typ, zone = _arrays(100, seed=3)
    path = tmp_path / "fluid_t=0.vtk"
    files.write_vtk(path, pos, vel, p, binary=binary)
    want = _reference_binary(pos, vel, p) if binary else _reference_ascii(pos, vel, p)
    assert path.read_bytes() == want
    # spot check of the big-endian encoding: 1.0f is 3f 80 00 00
    if binary:
        files.write_vtk(path, np.ones((1, 3)), np.zeros((1, 3)), np.ones(1), binary=True)
        assert b"POINTS 1 float\n\x3f\x80\x00\x00\x3f\x80\x00\x00\x3f\x80\x00\x00POINT_DATA" in path.read_bytes()


@pytest.mark.parametrize("binary", [False, True])
def test_vtk_read_back(tmp_path, lib, binary):
    pos, vel, acc, rho, p, typ, zone = _arrays(64, seed=4)
    path = tmp_path / "f.vtk"
    files.write_vtk(path, pos, vel, p, rho=rho, type=typ, binary=binary)
    d = files.read_vtk(path)
    assert d["binary"] == binary and d["title"] == "sph data"
    if binary:  # float32 words: exact
        np.testing.assert_array_equal(d["points"], pos)
        np.testing.assert_array_equal(d["velocity"], vel)
        np.testing.assert_array_equal(d["p"], p)
        np.testing.assert_array_equal(d["rho"], rho)
    else:  # 7 significant digits
        for k, want in (("points", pos), ("velocity", vel), ("p", p), ("rho", rho)):
            np.testing.assert_allclose(d[k], want.astype(np.float64), rtol=5.1e-7, atol=0, err_msg=k)
    np.testing.assert_array_equal(d["type"], typ)


def test_vtk_empty_and_errors(tmp_path, lib):
    e = np.zeros((0, 3), np.float32)
    files.write_vtk(tmp_path / "e.vtk", e, e, np.zeros(0, np.float32))
    d = files.read_vtk(tmp_path / "e.vtk")
    assert d["points"].shape == (0, 3) and d["p"].size == 0
    with pytest.raises(SphError, match="cannot create"):
        files.write_vtk(tmp_path / "no" / "such" / "dir.vtk", e, e, np.zeros(0, np.float32))


def test_restore_rejects_non_checkpoints(tmp_path, lib):
    """The checkpoint header and checksum are checked before any device work."""
    from tmp_sph_amd import sph
    bad = tmp_path / "x.ckpt"
    bad.write_bytes(b"not a checkpoint at all, just text")
    with pytest.raises(SphError, match="not a checkpoint"):
        sph.Simulation.restore(bad)
    with pytest.raises(SphError, match="cannot open"):
        sph.Simulation.restore(tmp_path / "missing.ckpt")
    # a valid magic with a truncated body
    bad.write_bytes(b"SPHCKPT1" + b"\0" * 4)
    with pytest.raises(SphError, match="not a checkpoint"):
        sph.Simulation.restore(bad)
