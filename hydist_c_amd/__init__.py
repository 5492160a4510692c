"""MI355X (gfx950) implementation of the Hydist_C ADI shallow-water and
suspended-sediment timestep, behind the reference's host surface.

The compute path is the HIP library libhydist_amd.so (C ABI in
include/hydist.h); this package holds its Python binding, the scenario
generators and the build recipe.
"""
from . import abi, scenarios  # noqa: F401

__all__ = ["abi", "scenarios", "engine", "build"]
