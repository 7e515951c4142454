"""Simulation1D: the 1-D profile container of the drop-in
(diffsims/simulations/simulation1d.py:31-73, without plotting)."""

from __future__ import annotations

import numpy as np

__all__ = ["Simulation1D"]


class Simulation1D:
    """Reflection families of one phase: |g| (1/A), intensities (maximum 100)
    and indices (hkil for hexagonal lattices)."""

    def __init__(self, phase, reciprocal_spacing, intensities, hkl, reciprocal_radius, wavelength):
        self.phase = phase
        self.reciprocal_spacing = reciprocal_spacing
        self.intensities = intensities
        self.hkl = hkl
        self.reciprocal_radius = reciprocal_radius
        self.wavelength = wavelength

    def __repr__(self):
        return f"Simulation1D(name: {self.phase.name}, wavelength: {self.wavelength})"

    @property
    def theta(self):
        return np.arctan2(np.array(self.reciprocal_spacing), 1 / self.wavelength)
