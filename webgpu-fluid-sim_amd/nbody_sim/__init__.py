"""nbody_sim -- MI355X Barnes-Hut step behind the reference's Simulation API.

Python host side of libnbody_sim.so (HIP kernels for gfx950 behind the C ABI
in include/nbody_sim.h).  ``Simulation`` and ``Config`` mirror
src/simulation.ts and src/config.ts of the reference.
"""
from ._lib import NbodyError, lib  # noqa: F401
from .config import Config, getDefaultConfig  # noqa: F401
from .simulation import SCENARIOS, Ensemble, Simulation, generate, generate_radius  # noqa: F401

__all__ = ["Config", "getDefaultConfig", "Simulation", "Ensemble", "generate", "generate_radius",
           "SCENARIOS",
           "NbodyError", "lib"]
