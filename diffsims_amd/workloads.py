"""Seeded synthetic workloads (phases, orientations, detector settings).

Each builder fixes every choice the workload description leaves open, so the
benchmark, the tests and the documentation use one definition:

* ``small_fcc_al``      FCC Al, 500 random orientations, 144x144 (CPU-runnable case)
* ``template_library``  orthorhombic 40-atom cell, 200,000-orientation SO(3) grid, 256x256
* ``refinement_monoclinic``  24-atom monoclinic cell, 4,096 perturbed targets (fwd+bwd)
* ``ragged_library``  64 triclinic phases of uneven size, ~10^6 orientations, 128x128
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


def haar_euler_rows(n: int, rng) -> np.ndarray:
    """Haar-uniform Bunge angles drawn row by row (three uniforms per row), so
    the first m rows do not depend on n."""
    u = rng.uniform(0, 1, (n, 3))
    return np.stack([2 * np.pi * u[:, 0], np.arccos(2 * u[:, 1] - 1), 2 * np.pi * u[:, 2]], axis=1)


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


@dataclass
class RefinementWorkload:
    """Differentiable refinement: a model phase/orientation set and the
    perturbed ground truth the targets are rendered from."""
    name: str
    phase: PhaseArrays              # the model's starting point
    euler: np.ndarray               # (n, 3) radians, model orientations
    truth_phase: PhaseArrays        # perturbed parameters that render the targets
    truth_euler: np.ndarray
    settings: SimSettings
    reciprocal_radius: float
    shape: tuple
    calibration: float
    sigma: float
    noise: float                    # Gaussian noise sigma as a fraction of each pattern's max
    seed: int


def refinement_monoclinic(n_patterns: int = 4096, seed: int = SEED + 3) -> RefinementWorkload:
    """24-atom monoclinic cell: a, b, c ~ U[5, 10] A, beta ~ U[95, 115] deg,
    fractional positions U[0, 1), Z in {8, 12, 14, 26} equiprobable,
    B ~ U[0.3, 1.0] A^2; r = 1.5 1/A, max_excitation_error 0.02, lorentzian,
    200 kV, no direct beam; n random orientations.  Targets: 256x256 sub-pixel
    patterns (0.015 1/A/px, sigma 1 px, clip 1, normalised) of the ground truth
    perturbed by orientation U[-0.5, 0.5] deg per Euler angle, lattice
    (a, b, c, beta) x (1 + U[-1 %, 1 %]), positions U[-0.02, 0.02], B
    U[-0.2, 0.2] (kept >= 0.05), plus Gaussian noise of 1 % of each pattern's
    maximum."""
    rng = np.random.default_rng(seed)
    abc = tuple(float(v) for v in rng.uniform(5, 10, 3))
    beta = float(rng.uniform(95, 115))
    Z = rng.choice([8, 12, 14, 26], 24)
    xyz = rng.uniform(0, 1, (24, 3))
    B = rng.uniform(0.3, 1.0, 24)
    symbols = [symbol(int(z)) for z in Z]
    model = PhaseArrays(abc, (90.0, beta, 90.0), symbols, xyz, np.ones(24), B, "mono24")
    euler = random_euler(n_patterns, rng)
    f = 1 + rng.uniform(-0.01, 0.01, 4)
    t_abc = tuple(float(v) for v in np.array(abc) * f[:3])
    truth = PhaseArrays(t_abc, (90.0, beta * float(f[3]), 90.0), symbols,
                        xyz + rng.uniform(-0.02, 0.02, xyz.shape), np.ones(24),
                        np.maximum(B + rng.uniform(-0.2, 0.2, 24), 0.05), "mono24-truth")
    t_euler = euler + np.deg2rad(rng.uniform(-0.5, 0.5, euler.shape))
    st = SimSettings(accelerating_voltage=200, max_excitation_error=0.02, shape_model="lorentzian",
                     minimum_intensity=1e-20)
    return RefinementWorkload("mono24-refinement", model, euler, truth, t_euler, st, 1.5,
                              (256, 256), 0.015, 1.0, 0.01, seed)


def ragged_library(n_phases: int = 64, orient_scale: float = 1.0, seed: int = SEED + 2) -> Workload:
    """Ragged multi-phase batch: per phase, atom count log-uniform in [1, 200],
    triclinic cell with edges U[3, 25] A and angles U[80, 110] deg, Z uniform
    in [3, 42], B U[0.2, 1.5] A^2, occupancy 1, a log-uniform 2,000-50,000
    Haar-random orientations (times ``orient_scale``; the phases do not depend
    on it and a reduced batch holds the first orientations of each phase's
    full set) and reciprocal_radius U[1.0, 2.5] 1/A.  200 kV,
    max_excitation_error 0.02, lorentzian, no direct beam; 128x128 f32
    patterns at 0.04 1/A/px (the frame holds r <= 2.56), sigma 1 px, fast
    render, normalised."""
    rng = np.random.default_rng(seed)
    phases, rots, radii = [], [], []
    for i in range(n_phases):
        n_atoms = int(round(np.exp(rng.uniform(np.log(1), np.log(200)))))
        abc = tuple(float(v) for v in rng.uniform(3, 25, 3))
        angles = tuple(float(v) for v in rng.uniform(80, 110, 3))
        Z = rng.integers(3, 43, n_atoms)
        xyz = rng.uniform(0, 1, (n_atoms, 3))
        B = rng.uniform(0.2, 1.5, n_atoms)
        phases.append(PhaseArrays(abc, angles, [symbol(int(z)) for z in Z], xyz, np.ones(n_atoms),
                                  B, f"ragged{i:02d}"))
        n_o = max(1, int(round(np.exp(rng.uniform(np.log(2000), np.log(50000))) * orient_scale)))
        radii.append(float(rng.uniform(1.0, 2.5)))
        # orientations from a stream of their own, one row at a time, so the
        # phases do not depend on orient_scale and a reduced batch holds the
        # first orientations of the full one
        rots.append(Rotation.from_euler(haar_euler_rows(n_o, np.random.default_rng([seed, i])))
                    .to_matrix())
    st = SimSettings(accelerating_voltage=200, max_excitation_error=0.02, shape_model="lorentzian",
                     minimum_intensity=1e-20)
    return Workload("ragged-64", phases, rots, radii, False, st, (128, 128), 0.04, 1.0,
                    notes={"orient_scale": orient_scale})


def large_precession(n_orient: int = 1_000_000, seed: int = SEED + 4):
    """Large triclinic cell with precession: 500 atoms, edges U[15, 30] A,
    angles U[80, 110] deg, Z uniform in [1, 30], B U[0.5, 2.0] A^2, occupancy 1;
    reciprocal_radius 3.0 1/A, precession 1 deg with the precession-integrated
    Lorentzian (approximate_precession), max_excitation_error 0.02, 200 kV, no
    direct beam; ``n_orient`` Haar-random orientations processed in chunks of
    65,536; 512x512 sub-pixel patterns (0.012 1/A/px: the frame holds r <= 3.07,
    sigma 1 px, clip 1, normalised).  The loss of a chunk is
    sum_o sum_px (P_o - T)^2 against one target T (the pattern of a held-out
    orientation), and each chunk yields the orientation and B gradients."""
    rng = np.random.default_rng(seed)
    n_atoms = 500
    abc = tuple(float(v) for v in rng.uniform(15, 30, 3))
    angles = tuple(float(v) for v in rng.uniform(80, 110, 3))
    Z = rng.integers(1, 31, n_atoms)
    pa = PhaseArrays(abc, angles, [symbol(int(z)) for z in Z], rng.uniform(0, 1, (n_atoms, 3)),
                     np.ones(n_atoms), rng.uniform(0.5, 2.0, n_atoms), "tri500")
    euler = random_euler(n_orient, rng)
    target_euler = random_euler(1, rng)
    st = SimSettings(accelerating_voltage=200, max_excitation_error=0.02, shape_model="lorentzian",
                     precession_angle=1.0, approximate_precession=True, minimum_intensity=1e-20)
    return {"phase": pa, "euler": euler, "target_euler": target_euler, "settings": st,
            "reciprocal_radius": 3.0, "shape": (512, 512), "calibration": 0.012, "sigma": 1.0,
            "clip": 1.0, "chunk": 65536}
