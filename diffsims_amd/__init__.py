"""diffsims_amd — MI355X-native kinematical electron-diffraction pattern simulation.

Drop-in for diffsims' ``SimulationGenerator.calculate_diffraction2d`` path:
the per-phase reflection tables, the per-orientation Ewald-sphere selection and
the pattern rendering run in hand-written HIP kernels (gfx950) behind the C ABI
of ``include/diffsims_amd.h``.  Importing the package does not need a GPU; the
compute entry points raise when the native library or the GPU is missing.
"""

from .crystallography import Atom, Lattice, Phase, Structure
from .diffracting_vector import DiffractingVector
from .rotation import Rotation, uniform_so3_grid
from .simulation1d import Simulation1D
from .simulation2d import Simulation2D, get_closest
from .simulation_generator import SimulationGenerator
from .legacy import (DiffractionGenerator, DiffractionLibrary, DiffractionLibraryGenerator,
                     DiffractionSimulation, ProfileSimulation, StructureLibrary)
from . import shape_factor_models

__version__ = "0.1.0"

__all__ = [
    "Atom",
    "DiffractingVector",
    "DiffractionGenerator",
    "DiffractionLibrary",
    "DiffractionLibraryGenerator",
    "DiffractionSimulation",
    "Lattice",
    "Phase",
    "ProfileSimulation",
    "Rotation",
    "Simulation1D",
    "Simulation2D",
    "SimulationGenerator",
    "Structure",
    "StructureLibrary",
    "get_closest",
    "shape_factor_models",
    "uniform_so3_grid",
]
