"""On-disk formats of the reference's I/O layer (src/io/io.cu), read and written
by the host C++ code of the library.

* :func:`read_dat` / :func:`write_dat` -- the initial-state directory of
  ``ini_fluid`` (src/io/io.cu:6-111): ``h_pos.dat``, ``h_vel.dat``,
  ``h_ac.dat``, ``rhop.dat``, ``p.dat``, ``particle_type.dat``,
  ``particle_zone.dat`` as whitespace-separated text.
* :func:`write_vtk` -- a legacy-VTK frame of ``output_fluid_file``: ASCII as
  src/io/io.cu:164-196 prints it, binary (big-endian float32) as
  opt_src_v3/io/io.cu:121-200 writes it.
* :func:`read_vtk` -- a reader for those frames (both encodings; binary blocks
  with or without a trailing newline).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import call, load

__all__ = ["count_dat", "read_dat", "write_dat", "write_vtk", "read_vtk"]


def _hp(a):
    return C.c_void_p(0) if a is None else a.ctypes.data_as(C.c_void_p)


def _b(path) -> bytes:
    return os.fsencode(os.fspath(path))


def count_dat(directory) -> int:
    """Particle count of an initial-state directory (values in rhop.dat)."""
    load()
    n = C.c_int()
    call("sph_count_dat", _b(directory), C.byref(n))
    return n.value


def read_dat(directory, n: int | None = None, *, acc: bool = True, zone: bool = True) -> dict:
    """Reads the seven initial-state files; ``n`` defaults to :func:`count_dat`."""
    load()
    n = count_dat(directory) if n is None else int(n)
    out = {"pos": np.empty((n, 3), np.float32), "vel": np.empty((n, 3), np.float32),
           "acc": np.empty((n, 3), np.float32) if acc else None, "rho": np.empty(n, np.float32),
           "p": np.empty(n, np.float32), "type": np.empty(n, np.int32),
           "zone": np.empty(n, np.int32) if zone else None}
    call("sph_read_dat", _b(directory), n, *(_hp(out[k]) for k in ("pos", "vel", "acc", "rho", "p", "type", "zone")))
    return {k: v for k, v in out.items() if v is not None}


def write_dat(directory, pos, vel, rho, p, type, acc=None, zone=None) -> None:
    load()
    os.makedirs(directory, exist_ok=True)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel, np.float32).reshape(n, 3)
    acc = None if acc is None else np.ascontiguousarray(acc, np.float32).reshape(n, 3)
    rho = np.ascontiguousarray(rho, np.float32).reshape(n)
    p = np.ascontiguousarray(p, np.float32).reshape(n)
    type = np.ascontiguousarray(type, np.int32).reshape(n)
    zone = None if zone is None else np.ascontiguousarray(zone, np.int32).reshape(n)
    call("sph_write_dat", _b(directory), n, _hp(pos), _hp(vel), _hp(acc), _hp(rho), _hp(p), _hp(type), _hp(zone))


def write_vtk(path, pos, vel, p, rho=None, type=None, binary: bool = False) -> None:
    load()
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel, np.float32).reshape(n, 3)
    p = np.ascontiguousarray(p, np.float32).reshape(n)
    rho = None if rho is None else np.ascontiguousarray(rho, np.float32).reshape(n)
    type = None if type is None else np.ascontiguousarray(type, np.int32).reshape(n)
    call("sph_write_vtk", _b(path), n, _hp(pos), _hp(vel), _hp(p), _hp(rho), _hp(type), int(bool(binary)))


def read_vtk(path) -> dict:
    """Parses a frame written by :func:`write_vtk` (or by the reference):
    returns ``points`` (n, 3), each SCALARS section by name and each VECTORS
    section by name, as float64 (ints for int sections), plus ``binary``."""
    data = open(path, "rb").read()
    pos = 0

    def line() -> str:
        nonlocal pos
        while pos < len(data) and data[pos] in b" \t\r\n":
            pos += 1
        e = data.index(b"\n", pos)
        s = data[pos:e].decode("ascii").strip()
        pos = e + 1
        return s

    def values(count: int, kind: str):
        nonlocal pos
        if binary:
            dt = ">i4" if kind == "int" else ">f4"
            v = np.frombuffer(data, dt, count, pos)
            pos += 4 * count
            if pos < len(data) and data[pos:pos + 1] == b"\n":
                pos += 1
            return v.astype(np.int64 if kind == "int" else np.float64)
        toks = []
        while len(toks) < count:
            toks += line().split()
        if len(toks) != count:
            raise ValueError("value lines do not match the section size")
        return np.array(toks, dtype=np.int64 if kind == "int" else np.float64)

    out: dict = {}
    if line() != "# vtk DataFile Version 3.0":
        raise ValueError("not a legacy VTK file")
    out["title"] = line()
    enc = line()
    if enc not in ("ASCII", "BINARY"):
        raise ValueError(f"unknown encoding {enc!r}")
    binary = enc == "BINARY"
    out["binary"] = binary
    if line() != "DATASET UNSTRUCTURED_GRID":
        raise ValueError("expected an unstructured grid")
    n = 0
    while True:
        while pos < len(data) and data[pos] in b" \t\r\n":
            pos += 1
        if pos >= len(data):
            break
        head = line().split()
        if head[0] == "POINTS":
            n = int(head[1])
            out["points"] = values(3 * n, head[2]).reshape(n, 3)
        elif head[0] == "POINT_DATA":
            if int(head[1]) != n:
                raise ValueError("POINT_DATA size differs from POINTS")
        elif head[0] == "SCALARS":
            if line() != "LOOKUP_TABLE DEFAULT":
                raise ValueError("expected LOOKUP_TABLE DEFAULT")
            out[head[1]] = values(n, head[2])
        elif head[0] == "VECTORS":
            out[head[1]] = values(3 * n, head[2]).reshape(n, 3)
        else:
            raise ValueError(f"unknown section {head[0]!r}")
    return out
