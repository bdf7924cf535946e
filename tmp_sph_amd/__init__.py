"""MI355X-native delta-SPH dam-break step (hand-written HIP for gfx950).

``tmp_sph_amd.sph`` holds the reference call surface and the fast-path
:class:`~tmp_sph_amd.sph.Simulation`; the compute runs in
``libsph_mi355x.so`` (built by ``tmp_sph_amd.build``).
"""
from . import sph  # noqa: F401
from .sph import Simulation, dambreak_cfg, make_case, reference_params  # noqa: F401

__all__ = ["sph", "Simulation", "dambreak_cfg", "make_case", "reference_params"]
