"""Simulation2D: result container of calculate_diffraction2d.

Mirrors diffsims/simulations/simulation2d.py:44-474 (nesting, slicing with
``iphase``/``irot``, iteration, ``get_simulation``, coordinate transforms,
``polar_flatten_simulations`` and ``get_diffraction_pattern``).  Rendering runs
on the GPU; ``get_diffraction_patterns`` renders every simulation of the
object in one launch per phase.  Interactive plotting (matplotlib widgets) is
not provided.
"""

from __future__ import annotations

import copy
from typing import Sequence, Tuple

import numpy as np

from .diffracting_vector import DiffractingVector

__all__ = ["Simulation2D", "get_closest"]


def _is_rotation(obj) -> bool:
    return hasattr(obj, "to_matrix") and hasattr(obj, "size") and not isinstance(obj, (list, tuple, np.ndarray))


def _is_phase(obj) -> bool:
    return hasattr(obj, "structure") and not isinstance(obj, (list, tuple, np.ndarray))


class PhaseGetter:
    """``sim.iphase[i]`` / ``sim.iphase["name"]`` (simulation2d.py:44-73)."""

    def __init__(self, simulation):
        self.simulation = simulation

    def __getitem__(self, item):
        all_phases = self.simulation.phases
        if _is_phase(all_phases):
            raise ValueError("Only one phase in the simulation")
        elif isinstance(item, str):
            ind = [phase.name for phase in all_phases].index(item)
        elif isinstance(item, (int, np.integer, slice)):
            ind = item
        else:
            raise ValueError("Item must be a string or integer")
        return Simulation2D(phases=all_phases[ind], coordinates=self.simulation.coordinates[ind],
                            rotations=self.simulation.rotations[ind],
                            simulation_generator=self.simulation.simulation_generator)


class RotationGetter:
    """``sim.irot[i]`` / slices (simulation2d.py:76-106)."""

    def __init__(self, simulation):
        self.simulation = simulation

    def __getitem__(self, item):
        sim = self.simulation
        if sim.current_size == 1:
            raise ValueError("Only one rotation in the simulation")
        elif _is_phase(sim.phases):
            coords = sim.coordinates[item]
            phases = sim.phases
            rotations = sim.rotations[item]
        else:
            coords = [c[item] for c in sim.coordinates]
            phases = sim.phases
            rotations = [rot[item] for rot in sim.rotations]
        return Simulation2D(phases=phases, coordinates=coords, rotations=rotations,
                            simulation_generator=sim.simulation_generator)


class Simulation2D:
    """Simulated spot lists for one or more phases and orientations."""

    def __init__(self, phases, coordinates, rotations, simulation_generator,
                 reciprocal_radius: float = 1.0):
        if _is_rotation(rotations) and rotations.size == 1:
            if not isinstance(coordinates, DiffractingVector):
                raise ValueError("If there is only one rotation, then the coordinates must be a "
                                 "DiffractingVector object")
        elif _is_rotation(rotations):
            coordinates = _object_array(coordinates)
            if coordinates.size != rotations.size:
                raise ValueError(f"The number of rotations: {rotations.size} must match the number "
                                 f"of coordinates {coordinates.size}")
        else:  # iterable of rotations, one per phase
            rotations = _object_array(rotations)
            coordinates = _object_array(coordinates)
            phases = _object_array(phases)
            if rotations.size != phases.size:
                raise ValueError(f"The number of rotations: {rotations.size} must match the number "
                                 f"of phases {phases.size}")
            for r, c in zip(rotations, coordinates):
                if isinstance(c, DiffractingVector):
                    c = np.array([c], dtype=object)
                if r.size != len(c):
                    raise ValueError(f"The number of rotations: {r.size} must match the number of "
                                     f"coordinates {len(c)}")
        self.phases = phases
        self.rotations = rotations
        self.coordinates = coordinates
        self.simulation_generator = simulation_generator
        self.phase_index = 0
        self.rotation_index = 0
        self.reciporical_radius = reciprocal_radius  # [sic], the reference's attribute name
        self.iphase = PhaseGetter(self)
        self.irot = RotationGetter(self)

    # -- indexing ---------------------------------------------------------------
    def get_simulation(self, item):
        """(rotation, phase index, coordinates) of flat simulation ``item``."""
        if self.has_multiple_phases:
            cumsum = np.cumsum(self._num_rotations())
            ind = int(np.searchsorted(cumsum, item, side="right"))
            cumsum = np.insert(cumsum, 0, 0)
            num_rot = cumsum[ind]
            if self.has_multiple_rotations[ind]:
                return (self.rotations[ind][item - num_rot], ind,
                        self.coordinates[ind][item - num_rot])
            return self.rotations[ind], ind, self.coordinates[ind]
        elif self.has_multiple_rotations:
            return self.rotations[item], 0, self.coordinates[item]
        return self.rotations[item], 0, self.coordinates

    def _num_rotations(self):
        if self.has_multiple_phases:
            return [r.size for r in self.rotations]
        return self.rotations.size

    def __iter__(self):
        return self

    def __next__(self):
        if self.phase_index == self.num_phases:
            self.phase_index = 0
            raise StopIteration
        coords = self.coordinates[self.phase_index] if self.has_multiple_phases else self.coordinates
        multi = (self.has_multiple_rotations[self.phase_index] if self.has_multiple_phases
                 else self.has_multiple_rotations)
        if multi:
            coords = coords[self.rotation_index]
        if self.rotation_index + 1 == self.current_size:
            self.rotation_index = 0
            self.phase_index += 1
        else:
            self.rotation_index += 1
        return coords

    @property
    def current_size(self) -> int:
        if self.has_multiple_phases:
            return self.rotations[self.phase_index].size
        return self.rotations.size

    def deepcopy(self):
        return copy.deepcopy(self)

    @property
    def num_phases(self) -> int:
        return len(self.phases) if hasattr(self.phases, "__len__") and not _is_phase(self.phases) else 1

    @property
    def has_multiple_phases(self) -> bool:
        return self.num_phases > 1

    @property
    def has_multiple_rotations(self):
        if _is_rotation(self.rotations):
            return self.rotations.size > 1
        return [r.size > 1 for r in self.rotations]

    @property
    def current_phase(self):
        return self.phases[self.phase_index] if self.has_multiple_phases else self.phases

    def get_current_coordinates(self) -> DiffractingVector:
        if self.has_multiple_phases:
            c = self.coordinates[self.phase_index]
            multi = self.has_multiple_rotations[self.phase_index]
            return copy.deepcopy(c[self.rotation_index] if multi else c)
        if self.has_multiple_rotations:
            return copy.deepcopy(self.coordinates[self.rotation_index])
        return copy.deepcopy(self.coordinates)

    def get_current_rotation_matrix(self) -> np.ndarray:
        if self.has_multiple_phases:
            return copy.deepcopy(self.rotations[self.phase_index].to_matrix()[self.rotation_index])
        return copy.deepcopy(self.rotations.to_matrix()[self.rotation_index])

    # -- coordinates ----------------------------------------------------------------
    def _get_transformed_coordinates(self, angle: float, center: Sequence = (0, 0),
                                     mirrored: bool = False, units: str = "real",
                                     calibration: float = None) -> DiffractingVector:
        """Scale, rotate in-plane, mirror and shift (simulation2d.py:261-285); the
        assignments are in place, so integer coordinates stay integer."""
        coords = self.get_current_coordinates()
        if units != "real":
            center = np.array(center)
            coords.data = coords.data / calibration
        cx, cy = center
        x = coords.data[:, 0]
        y = coords.data[:, 1]
        mf = -1 if mirrored else 1
        theta = mf * np.arctan2(y, x) + np.deg2rad(angle)
        rd = np.sqrt(x ** 2 + y ** 2)
        coords[:, 0] = rd * np.cos(theta) + cx
        coords[:, 1] = rd * np.sin(theta) + cy
        return coords

    def rotate_shift_coordinates(self, angle: float, center: Sequence = (0, 0),
                                 mirrored: bool = False) -> DiffractingVector:
        return self._get_transformed_coordinates(angle, center, mirrored, units="real")

    def polar_flatten_simulations(self, radial_axes=None, azimuthal_axes=None):
        """(r, theta, intensity) templates of shape (n_simulations, max spots)
        (simulation2d.py:313-355)."""
        flattened = [sim for sim in self]
        max_num_spots = max(v.size for v in flattened)
        n = len(flattened)
        r_t = np.zeros((n, max_num_spots))
        t_t = np.zeros((n, max_num_spots))
        i_t = np.zeros((n, max_num_spots))
        for i, v in enumerate(flattened):
            r, t = v.to_flat_polar()
            inten = v.intensity
            if radial_axes is not None and azimuthal_axes is not None:
                r = get_closest(radial_axes, r)
                t = get_closest(azimuthal_axes, t)
                mask = (r < len(radial_axes) - 1) & (t < len(azimuthal_axes) - 1)
                r, t, inten = r[mask], t[mask], inten[mask]
            r_t[i, :len(r)] = r
            t_t[i, :len(r)] = t
            i_t[i, :len(inten)] = inten
        if radial_axes is not None and azimuthal_axes is not None:
            r_t = np.array(r_t, dtype=int)
            t_t = np.array(t_t, dtype=int)
        return r_t, t_t, i_t

    # -- rendering ------------------------------------------------------------------
    def get_diffraction_pattern(self, shape: Tuple[int, int] = None, sigma: float = 10,
                                direct_beam_position: Tuple[int, int] = None,
                                in_plane_angle: float = 0, calibration: float = 0.01,
                                mirrored: bool = False, fast: bool = True, normalize: bool = True,
                                clip_threshold: float = 1) -> np.ndarray:
        """The current pattern rendered on the GPU (simulation2d.py:357-442).

        The coordinate transform runs on the host with the reference's dtype
        semantics; the splat, Gaussian and normalisation run in the HIP renderer.
        """
        from . import engine
        import torch

        coords_dtype = self.get_current_coordinates().data.dtype
        integer_coords = np.issubdtype(coords_dtype, np.integer)
        if direct_beam_position is None:
            if fast or integer_coords:
                direct_beam_position = (shape[1] // 2, shape[0] // 2)
            else:
                direct_beam_position = ((shape[1] - 1) / 2, (shape[0] - 1) / 2)
        tr = self._get_transformed_coordinates(in_plane_angle, direct_beam_position, mirrored,
                                               units="pixel", calibration=calibration)
        d = tr.data
        in_frame = (d[:, 0] >= 0) & (d[:, 0] < shape[1]) & (d[:, 1] >= 0) & (d[:, 1] < shape[0])
        pix = d[in_frame][:, :2]
        inten = np.asarray(tr.intensity, dtype=np.float64)[in_frame]
        if inten.shape[0] == 0:
            return np.zeros(shape)
        integer_path = fast or np.issubdtype(pix.dtype, np.integer)
        if integer_path:
            pix = pix.astype(int)
        dev = engine._device()
        spots = np.zeros((1, pix.shape[0], 4), dtype=np.float32)
        spots[0, :, 0] = pix[:, 0]
        spots[0, :, 1] = pix[:, 1]
        spots[0, :, 3] = inten
        t_spots = torch.from_numpy(spots).to(dev)
        t_count = torch.tensor([pix.shape[0]], dtype=torch.int32, device=dev)
        out = engine.render_spots(t_spots, t_count, pix.shape[0], shape, sigma, 1.0,
                                  direct_beam_position=(0.0, 0.0), normalize=normalize,
                                  fast=integer_path, clip_threshold=clip_threshold)
        return out[0].cpu().numpy().astype(np.float64)

    def get_diffraction_patterns(self, shape: Tuple[int, int], sigma: float = 10,
                                 direct_beam_position=None, in_plane_angle: float = 0,
                                 calibration: float = 0.01, mirrored: bool = False,
                                 fast: bool = True, normalize: bool = True,
                                 clip_threshold: float = 1, as_numpy: bool = False):
        """Every simulation of this object rendered on the GPU in flat
        ``get_simulation`` order: a float32 tensor [n, H, W] (or NumPy array).

        Extension of the reference API; spot coordinates are transformed in the
        kernel in FP32.
        """
        from . import engine
        import torch

        flat = [v for v in self]
        dev = engine._device()
        cap = max(1, max(v.size for v in flat))
        spots = np.zeros((len(flat), cap, 4), dtype=np.float32)
        count = np.zeros(len(flat), dtype=np.int32)
        for i, v in enumerate(flat):
            spots[i, :v.size, :3] = np.asarray(v.data, dtype=np.float32)
            spots[i, :v.size, 3] = np.asarray(v.intensity, dtype=np.float32)
            count[i] = v.size
        out = engine.render_spots(torch.from_numpy(spots).to(dev), torch.from_numpy(count).to(dev),
                                  cap, shape, sigma, calibration, direct_beam_position,
                                  in_plane_angle, mirrored, normalize, fast, clip_threshold)
        return out.cpu().numpy() if as_numpy else out

    def plot(self, *args, **kwargs):
        raise NotImplementedError("interactive plotting is not part of diffsims_amd")


def _object_array(items) -> np.ndarray:
    if isinstance(items, np.ndarray) and items.dtype == object:
        return items
    arr = np.empty(len(items), dtype=object)
    for i, it in enumerate(items):
        arr[i] = it
    return arr


def get_closest(array, values):
    """Index of the nearest entry of a sorted array for each value
    (simulation2d.py:767-781)."""
    array = np.array(array)
    idxs = np.searchsorted(array, values, side="left")
    prev_idx_is_less = (idxs == len(array)) | (
        np.fabs(values - array[np.maximum(idxs - 1, 0)])
        < np.fabs(values - array[np.minimum(idxs, len(array) - 1)]))
    idxs[prev_idx_is_less] -= 1
    return idxs
