"""Seeded synthetic workloads (phases, orientations, detector settings).

Each builder fixes every choice the workload description leaves open, so the
benchmark, the tests and the documentation use one definition:

* ``small_fcc_al``      FCC Al, 500 random orientations, 144x144 (CPU-runnable case)
* ``template_library``  orthorhombic 40-atom cell, 200,000-orientation SO(3) grid, 256x256
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from .crystallography import PhaseArrays
from .engine import SimSettings
from .rotation import Rotation, uniform_so3_grid
from .scattering import symbol

SEED = 20261202


@dataclass
class Workload:
    name: str
    phases: List[PhaseArrays]
    rotations: List[np.ndarray]        # per phase: (n, 3, 3) orix to_matrix()
    reciprocal_radius: List[float]
    with_direct_beam: bool
    settings: SimSettings
    shape: tuple
    calibration: float
    sigma: float
    notes: dict = field(default_factory=dict)

    @property
    def n_patterns(self) -> int:
        return int(sum(r.shape[0] for r in self.rotations))


def random_euler(n: int, rng) -> np.ndarray:
    """Haar-uniform Bunge angles: phi1, phi2 ~ U[0, 2pi), Phi = arccos U[-1, 1]."""
    p1 = rng.uniform(0, 2 * np.pi, n)
    P = np.arccos(rng.uniform(-1, 1, n))
    p2 = rng.uniform(0, 2 * np.pi, n)
    return np.stack([p1, P, p2], axis=1)


def small_fcc_al(n_orient: int = 500) -> Workload:
    """FCC Al a = 4.05 A, 4 atoms, B = 0.5 A^2; 200 kV, r = 1.0 1/A,
    max_excitation_error 0.01, linear, no direct beam; 144x144 at 0.01 1/A/px,
    sigma 1 px; seeded uniform-random orientations."""
    rng = np.random.default_rng(SEED)
    pa = PhaseArrays((4.05,) * 3, (90.0,) * 3, ["Al"] * 4,
                     np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], float),
                     np.ones(4), np.full(4, 0.5), "Al")
    M = Rotation.from_euler(random_euler(n_orient, rng)).to_matrix()
    st = SimSettings(accelerating_voltage=200, max_excitation_error=0.01, shape_model="linear",
                     minimum_intensity=1e-20)
    return Workload("fcc-al-500", [pa], [M], [1.0], False, st, (144, 144), 0.01, 1.0)


def ortho40_phase(seed: int = SEED + 1) -> PhaseArrays:
    """Orthorhombic cell, a, b, c ~ U[4, 12] A, 40 atoms at uniform fractional
    positions, Z in {6, 8, 14, 26} equiprobable, B ~ U[0.3, 1.0] A^2, occupancy 1."""
    rng = np.random.default_rng(seed)
    abc = tuple(float(v) for v in rng.uniform(4, 12, 3))
    Z = rng.choice([6, 8, 14, 26], 40)
    xyz = rng.uniform(0, 1, (40, 3))
    B = rng.uniform(0.3, 1.0, 40)
    return PhaseArrays(abc, (90.0,) * 3, [symbol(int(z)) for z in Z], xyz, np.ones(40), B, "ortho40")


def template_library(n_orient: int = 200_000) -> Workload:
    """The large single-phase template library: ortho40 cell, r = 2.0 1/A,
    max_excitation_error 0.02, lorentzian, 200 kV, no direct beam; n_orient
    rotations of a super-Fibonacci SO(3) grid; 256x256 f32 at 0.015 1/A/px,
    sigma 1 px."""
    pa = ortho40_phase()
    M = uniform_so3_grid(n_orient).to_matrix()
    st = SimSettings(accelerating_voltage=200, max_excitation_error=0.02, shape_model="lorentzian",
                     minimum_intensity=1e-20)
    return Workload("ortho40-library", [pa], [M], [2.0], False, st, (256, 256), 0.015, 1.0,
                    notes={"grid": "super-Fibonacci SO(3) spiral", "sigma_px": 1.0,
                           "with_direct_beam": False})
