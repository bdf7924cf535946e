"""Case builders: the dam-break configurations the benchmarks and tests use.

* ``c1`` .. ``c3``: the 2-D (dx = 0.01 and 0.001 m) and 3-D (dx = 0.005 m)
  dam breaks.
* :func:`sweep_cases`: the batched sweep -- uneven 2-D dam breaks whose column
  height H ~ U[0.2, 0.6] m, width/height ratio ~ U[0.5, 2] and resolution
  dx in {0.002, 0.004} m are drawn per instance from splitmix64(seed, k).
"""
from __future__ import annotations

from dataclasses import dataclass

from .sph import Case, dambreak_cfg, make_case

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """One output of the splitmix64 generator (Steele, Lea & Flood 2014) for state x."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def _uniforms(seed: int, k: int, count: int) -> list[float]:
    state = splitmix64((seed & MASK64) ^ splitmix64(k))
    out = []
    for _ in range(count):
        state = splitmix64(state)
        out.append((state >> 11) * (1.0 / (1 << 53)))
    return out


@dataclass
class SweepMember:
    k: int
    height: float
    width: float
    dx: float


def sweep_members(count: int = 64, seed: int = 20240607) -> list[SweepMember]:
    members = []
    for k in range(count):
        u = _uniforms(seed, k, 3)
        h = 0.2 + 0.4 * u[0]
        ratio = 0.5 + 1.5 * u[1]
        dx = 0.002 if u[2] < 0.5 else 0.004
        # column sizes on the lattice
        h = max(round(h / dx), 1) * dx
        w = max(round(h * ratio / dx), 1) * dx
        members.append(SweepMember(k, h, w, dx))
    return members


def sweep_cases(count: int = 64, seed: int = 20240607, tank=(4.0, 1.0, 1.0)) -> list[Case]:
    """Uneven 2-D dam breaks in a 4.0 x 1.0 m tank, one Case each."""
    cases = []
    for m in sweep_members(count, seed):
        cases.append(make_case(dambreak_cfg(dim=2, dx=m.dx, tank=tank, column=(m.width, 1.0, m.height))))
    return cases


def named(name: str):
    """The named configuration: (description, DamBreakCfg)."""
    table = {
        "c1": ("2-D dam break, 1.0x0.5 m column in 4.0x1.0 m tank, dx=0.01 m", dambreak_cfg(dim=2, dx=0.01)),
        "c2": ("2-D dam break, 1.0x0.5 m column in 4.0x1.0 m tank, dx=0.001 m", dambreak_cfg(dim=2, dx=0.001)),
        "c3": ("3-D dam break, 1.0x0.5x0.5 m column in 3.2x1.0x1.0 m tank, dx=0.005 m",
               dambreak_cfg(dim=3, dx=0.005, tank=(3.2, 1.0, 1.0), column=(1.0, 0.5, 0.5))),
    }
    return table[name]
