"""The legacy generator API on the GPU path.

Drop-ins for diffsims' older entry points, which simulate from a bare diffpy
``Structure`` in the structure's own Cartesian frame:

* :class:`DiffractionGenerator` -- ``calculate_ed_data`` (one orientation,
  rzxz Euler angles in degrees) and ``calculate_profile_data`` (1-D profile)
  (diffsims/generators/diffraction_generator.py:89-417);
* :class:`DiffractionLibraryGenerator` -- ``get_diffraction_library`` over a
  :class:`StructureLibrary` (diffsims/generators/library_generator.py:40-153),
  one reflection table per phase and one spot-kernel launch for all of its
  orientations;
* the result containers :class:`DiffractionSimulation`,
  :class:`ProfileSimulation` (diffsims/sims/diffraction_simulation.py) and
  :class:`DiffractionLibrary` (diffsims/libraries/diffraction_library.py).

The physics is the one of the current path (excitation error, shape factor,
relative cut) with the legacy reflection set (``dsx_hkl_points_in_sphere``:
ascending (h, k, l), |g| < r, (000) included) and frame (diffpy's standard
base, a* || x, c || z; :func:`engine.diffpy_standard_base`).  The rotation is
transforms3d's ``euler2mat(ai, aj, ak, "rzxz")`` (rotating axes:
Rz(ai) Rx(aj) Rz(ak)), applied as g_lab = R g.  Plotting and mask helpers of
the old containers are not part of this path.
"""

from __future__ import annotations

import copy

import numpy as np

from . import engine
from .crystallography import PhaseArrays, element_of
from .simulation_generator import SimulationGenerator, _flat_spots

__all__ = ["DiffractionGenerator", "DiffractionSimulation", "ProfileSimulation",
           "DiffractionLibraryGenerator", "DiffractionLibrary", "StructureLibrary",
           "euler2mat_rzxz"]


def euler2mat_rzxz(angles_deg) -> np.ndarray:
    """Rotation matrices of rzxz Euler angles in degrees, (n, 3) -> (n, 3, 3):
    rotating axes z, x', z'' (transforms3d ``euler2mat(..., axes="rzxz")``),
    M = Rz(ai) Rx(aj) Rz(ak) with active elementary rotations."""
    a = np.deg2rad(np.asarray(angles_deg, dtype=np.float64).reshape(-1, 3))
    ci, cj, ck = np.cos(a[:, 0]), np.cos(a[:, 1]), np.cos(a[:, 2])
    si, sj, sk = np.sin(a[:, 0]), np.sin(a[:, 1]), np.sin(a[:, 2])
    n = a.shape[0]
    z, o = np.zeros(n), np.ones(n)
    Rz1 = np.stack([np.stack([ci, -si, z], -1), np.stack([si, ci, z], -1),
                    np.stack([z, z, o], -1)], -2)
    Rx = np.stack([np.stack([o, z, z], -1), np.stack([z, cj, -sj], -1),
                   np.stack([z, sj, cj], -1)], -2)
    Rz2 = np.stack([np.stack([ck, -sk, z], -1), np.stack([sk, ck, z], -1),
                    np.stack([z, z, o], -1)], -2)
    return Rz1 @ Rx @ Rz2


def structure_arrays(structure, debye_waller_factors=None, name="") -> PhaseArrays:
    """A diffpy-style structure (atoms with ``element``, ``xyz``,
    ``occupancy``; a ``lattice``) as :class:`PhaseArrays`; B from the dict
    ``debye_waller_factors`` by element (missing: 0), as
    get_vectorized_list_for_atomic_scattering_factors (sim_utils.py:165-224)."""
    lat = structure.lattice
    atoms = list(structure)
    symbols = [str(a.element) for a in atoms]
    xyz = np.array([np.asarray(a.xyz, dtype=np.float64) for a in atoms]).reshape(-1, 3)
    occ = np.array([float(a.occupancy) for a in atoms], dtype=np.float64)
    dwf = debye_waller_factors or {}
    B = np.array([float(dwf.get(element_of(s), 0)) for s in symbols])
    return PhaseArrays((lat.a, lat.b, lat.c), (lat.alpha, lat.beta, lat.gamma), symbols, xyz, occ,
                       B, name)


def is_lattice_hexagonal(latt) -> bool:
    """a == b, alpha == beta == 90, gamma == 120 (sim_utils.py:477-493)."""
    return latt.a == latt.b and latt.alpha == 90 and latt.beta == 90 and latt.gamma == 120


def unique_families(hkls):
    """Families of indices equal up to permutation and sign
    (get_unique_families, sim_utils.py:101-136): {representative:
    multiplicity} in order of first appearance, the representative being the
    largest member in lexicographic order."""
    members = {}
    for hkl in hkls:
        key = tuple(sorted(abs(v) for v in hkl))
        members.setdefault(key, []).append(list(hkl))
    return {tuple(sorted(v)[-1]): len(v) for v in members.values()}


class DiffractionSimulation:
    """Spots of one legacy simulation (diffsims/sims/diffraction_simulation.py:
    31-352): coordinates (n, 2 or 3) in 1/A, indices (n, 3), intensities (n,),
    calibration, offset; the direct beam (all-zero coordinates) is masked out
    unless ``with_direct_beam``."""

    def __init__(self, coordinates, indices=None, intensities=None, calibration=None,
                 offset=(0.0, 0.0), with_direct_beam=False):
        coordinates = np.asarray(coordinates)
        if coordinates.ndim == 1:
            coordinates = coordinates[None, :]
        if indices is None:
            indices = np.full((coordinates.shape[0], 3), np.nan)
        if intensities is None:
            intensities = np.full((coordinates.shape[0]), np.nan)
        indices = np.asarray(indices)
        intensities = np.asarray(intensities)
        if (coordinates.shape[0] == indices.shape[0] == intensities.shape[0]
                and coordinates.ndim == indices.ndim == 2 and intensities.ndim == 1):
            self._coordinates = coordinates
            self._indices = indices
            self._intensities = intensities
        else:
            raise ValueError("Coordinate, intensity, and indices lists must be of the correct and "
                             "matching shape.")
        self.calibration = calibration
        self.offset = np.array(offset)
        self.with_direct_beam = with_direct_beam

    def __len__(self):
        return self.coordinates.shape[0]

    @property
    def size(self):
        return self.__len__()

    def __getitem__(self, sliced):
        coords = self.coordinates[sliced]
        inds = self.indices[sliced]
        ints = self.intensities[sliced]
        if coords.ndim == 1:
            coords = coords[None, :]
            inds = inds[None, :]
            ints = ints[None]
        if coords.ndim > 2 or coords.shape[1] > 3 or coords.shape[1] < 2:
            raise ValueError(f"Invalid slice: {sliced}")
        return DiffractionSimulation(coords, indices=inds, intensities=ints,
                                     calibration=self.calibration, offset=self.offset,
                                     with_direct_beam=self.with_direct_beam)

    def deepcopy(self):
        return copy.deepcopy(self)

    def __add__(self, other):
        new = self.deepcopy()
        new.extend(other)
        return new

    def extend(self, other):
        """Append the spots of another simulation (the reference stores the
        concatenated intensities under an attribute nothing reads, so its
        ``+`` loses them; they are kept here)."""
        self._coordinates = np.concatenate([self._coordinates, other._coordinates], axis=0)
        self._indices = np.concatenate([self._indices, other._indices], axis=0)
        self._intensities = np.concatenate([self._intensities, other._intensities], axis=0)

    @property
    def direct_beam_mask(self):
        if self.with_direct_beam:
            return np.ones_like(self._intensities, dtype=bool)
        return np.any(self._coordinates, axis=1)

    @property
    def coordinates(self):
        return self._coordinates[self.direct_beam_mask]

    @coordinates.setter
    def coordinates(self, coordinates):
        self._coordinates[self.direct_beam_mask] = coordinates

    @property
    def indices(self):
        return self._indices[self.direct_beam_mask]

    @indices.setter
    def indices(self, indices):
        self._indices[self.direct_beam_mask] = indices

    @property
    def intensities(self):
        return self._intensities[self.direct_beam_mask]

    @intensities.setter
    def intensities(self, intensities):
        self._intensities[self.direct_beam_mask] = intensities

    @property
    def calibration(self):
        return self._calibration

    @calibration.setter
    def calibration(self, calibration):
        if calibration is None:
            pass
        elif np.all(np.equal(calibration, 0)):
            raise ValueError("`calibration` cannot be zero.")
        elif isinstance(calibration, (float, int)):
            calibration = np.array((calibration, calibration))
        elif len(calibration) == 2:
            calibration = np.array(calibration)
        else:
            raise ValueError("`calibration` must be a float or length-2tuple of floats.")
        self._calibration = calibration

    @property
    def calibrated_coordinates(self):
        """Coordinates in pixels: (x, y) plus the offset over the calibration."""
        if self.calibration is not None:
            return (self.coordinates[:, :2] + self.offset) / self.calibration
        raise Exception("Pixel calibration is not set!")

    def _get_transformed_coordinates(self, angle, center=(0, 0), mirrored=False, units="real"):
        c = self.coordinates.copy() if units == "real" else self.calibrated_coordinates.copy()
        x, y = c[:, 0], c[:, 1]
        theta = (-1 if mirrored else 1) * np.arctan2(y, x) + np.deg2rad(angle)
        rd = np.sqrt(x ** 2 + y ** 2)
        c[:, 0] = rd * np.cos(theta) + center[0]
        c[:, 1] = rd * np.sin(theta) + center[1]
        return c

    def rotate_shift_coordinates(self, angle, center=(0, 0), mirrored=False):
        """Rotate (degrees), mirror and shift the spots in-plane."""
        self.coordinates = self._get_transformed_coordinates(angle, center, mirrored, units="real")

    def get_diffraction_pattern(self, shape=(512, 512), sigma=10, direct_beam_position=None,
                                in_plane_angle=0, mirrored=False):
        """The spots on a detector (diffraction_simulation.py:281-338): pixel
        coordinates (calibrated, rotated, shifted to the direct beam,
        truncated; the last spot on a pixel wins), scipy's Gaussian filter
        and division by the maximum, rendered by the GPU's fast renderer.
        Square detectors (the reference indexes its buffer [x, y] and
        transposes it)."""
        import torch

        H, W = int(shape[0]), int(shape[1])
        if direct_beam_position is None:
            direct_beam_position = (W // 2, H // 2)
        xy = self.calibrated_coordinates
        I = self.intensities
        n = len(I)
        cap = max(n, 1)
        spots = torch.zeros((1, cap, 4), dtype=torch.float32)
        if n:
            spots[0, :n, 0] = torch.from_numpy(np.ascontiguousarray(xy[:, 0], dtype=np.float32))
            spots[0, :n, 1] = torch.from_numpy(np.ascontiguousarray(xy[:, 1], dtype=np.float32))
            spots[0, :n, 3] = torch.from_numpy(np.ascontiguousarray(I, dtype=np.float32))
        dev = engine._device()
        spots = spots.to(dev)
        count = torch.tensor([n], dtype=torch.int32, device=dev)
        out = engine.render_spots(spots, count, cap, (H, W), sigma, 1.0,
                                  direct_beam_position=direct_beam_position,
                                  in_plane_angle=in_plane_angle, mirrored=mirrored)
        return out[0].double().cpu().numpy()


class ProfileSimulation:
    """1-D profile: magnitudes, intensities (max 100) and hkl labels
    (diffraction_simulation.py:355-505, without plotting)."""

    def __init__(self, magnitudes, intensities, hkls):
        self.magnitudes = magnitudes
        self.intensities = intensities
        self.hkls = hkls


class DiffractionGenerator(SimulationGenerator):
    """Legacy kinematic generator on the GPU
    (diffsims/generators/diffraction_generator.py:89-332); constructor,
    attributes and errors as the reference."""

    def __init__(self, accelerating_voltage, scattering_params="lobato", precession_angle=0,
                 shape_factor_model="lorentzian", approximate_precession=True,
                 minimum_intensity=1e-20, **kwargs):
        super().__init__(accelerating_voltage, scattering_params, precession_angle,
                         shape_factor_model, approximate_precession, minimum_intensity, **kwargs)

    def __repr__(self):
        return (f"DiffractionGenerator(accelerating_voltage={self.accelerating_voltage}, "
                f"scattering_params={self.scattering_params})")

    # -- tables and orientation batches -------------------------------------
    def _table(self, structure, reciprocal_radius, debye_waller_factors):
        pa = structure_arrays(structure, debye_waller_factors)
        return engine.build_legacy_table(pa, reciprocal_radius, self.scattering_params)

    def _simulate_batch(self, tab, angles_deg, reciprocal_radius, max_excitation_error,
                        shape_factor_width):
        """Spots of every orientation (rzxz degrees) of one table: flat host
        arrays (row index, lab xyz in FP64, I, counts)."""
        M = euler2mat_rzxz(angles_deg)
        # the spot kernel takes rows of R with g_lab = R g; rotation_matrices_to_device
        # transposes its argument, so it is handed R^T
        Rt = engine.rotation_matrices_to_device(np.transpose(M, (0, 2, 1)))
        model = self._device_model()
        if model is not None:
            batch = engine.simulate(tab, Rt, self._settings(max_excitation_error, shape_factor_width),
                                    reciprocal_radius)
            return _flat_spots(batch, np.transpose(M, (0, 2, 1)))
        # callables without a HIP kernel: the GPU selects, the host evaluates
        settings = self._settings(max_excitation_error, shape_factor_width, model="binary",
                                  minimum_intensity=-1.0)
        batch = engine.simulate(tab, Rt, settings, reciprocal_radius)
        idx, xyz, F2, counts = _flat_spots(batch, np.transpose(M, (0, 2, 1)))
        k = 1.0 / self.wavelength
        r = np.sqrt(np.sum(xyz[:, :2] ** 2, axis=1))
        s = (-np.sqrt(k ** 2 - r ** 2) + k) - xyz[:, 2]
        width = shape_factor_width if shape_factor_width is not None else max_excitation_error
        if self.precession_angle == 0:
            sf = self.shape_factor_model(s, width, **self.shape_factor_kwargs)
        else:
            from .simulation_generator import _precession_integral
            sf = _precession_integral(s, r, np.deg2rad(self.precession_angle),
                                      self.shape_factor_model, width, **self.shape_factor_kwargs)
        I = np.asarray(sf, dtype=np.float64) * F2
        keep = np.zeros(len(I), dtype=bool)
        start = 0
        for c in counts:
            sl = slice(start, start + int(c))
            start += int(c)
            if c:
                keep[sl] = I[sl] > np.max(I[sl]) * self.minimum_intensity
        new_counts = np.array([int(keep[s0:s0 + int(c)].sum()) for s0, c in
                               zip(np.concatenate([[0], np.cumsum(counts)[:-1]]), counts)],
                              dtype=np.int64)
        return idx[keep], xyz[keep], I[keep], new_counts

    def calculate_ed_data(self, structure, reciprocal_radius, rotation=(0, 0, 0),
                          with_direct_beam=True, max_excitation_error=1e-2,
                          shape_factor_width=None, debye_waller_factors={}):
        """Spots of ``structure`` at one orientation (rzxz Euler angles in
        degrees) (diffraction_generator.py:192-331)."""
        tab = self._table(structure, reciprocal_radius, debye_waller_factors)
        idx, xyz, I, counts = self._simulate_batch(tab, [rotation], reciprocal_radius,
                                                   max_excitation_error, shape_factor_width)
        hkl = tab.hkl_host().astype(np.float64)[idx]
        return DiffractionSimulation(coordinates=xyz, indices=hkl, intensities=I,
                                     with_direct_beam=with_direct_beam)

    def calculate_profile_data(self, structure, reciprocal_radius=1.0, minimum_intensity=1e-3,
                               debye_waller_factors={}):
        """1-D profile (diffraction_generator.py:334-417): the families of
        get_intensities_params (indices equal up to permutation and sign),
        intensity |F|^2 x multiplicity of each family's representative,
        labels from the indices (for hexagonal lattices the reference's
        Miller-Bravais step, kept as it is), scaled to a maximum of 100,
        families at or below ``minimum_intensity`` and (000) dropped."""
        tab = self._table(structure, reciprocal_radius, debye_waller_factors)
        hkl_all = tab.hkl_host()
        fam = unique_families(hkl_all.tolist())
        reps = np.array(list(fam), dtype=np.int64).reshape(-1, 3)
        mult = np.fromiter(fam.values(), dtype=float)
        pa = structure_arrays(structure, debye_waller_factors)
        ftab = engine.structure_factor_table(pa, reps, tab.rec, self.scattering_params,
                                             warn_missing=False)
        F2 = ftab.F2.double().cpu().numpy()
        # |g| of the representatives in FP64 (the reference's lattice distances)
        g_reps = np.sqrt(np.sum((reps.astype(np.float64) @ tab.rec.T) ** 2, axis=1)).tolist()
        i_hkl = F2 * mult
        if is_lattice_hexagonal(structure.lattice):
            # as the reference writes it: a 4-tuple of the first three index
            # rows (row0, row1, -row0 - row1, row2), so only four labelled
            # peaks remain (diffraction_generator.py:384-391)
            labels_idx = (reps[0], reps[1], -reps[0] - reps[1], reps[2])
        else:
            labels_idx = reps.tolist()
        labels = ["".join(str(int(x)) for x in xs) for xs in labels_idx]
        peaks = {}
        for lab, i, g in zip(labels, i_hkl, g_reps):
            peaks[lab] = [i, g]
        max_intensity = max(v[0] for v in peaks.values())
        x, y, hkls = [], [], []
        for k_, v in peaks.items():
            if v[0] / max_intensity * 100 > minimum_intensity and k_ != "000":
                x.append(v[1])
                y.append(v[0])
                hkls.append(k_)
        y = np.asarray(y) / max(y) * 100
        return ProfileSimulation(x, y, hkls)


class StructureLibrary:
    """Structures and their orientation lists (rzxz degrees)
    (diffsims/libraries/structure_library.py:25-114)."""

    def __init__(self, identifiers, structures, orientations):
        if len(identifiers) != len(structures):
            raise ValueError("Number of identifiers ({}) and structures ({}) must be the same."
                             .format(len(identifiers), len(structures)))
        if len(identifiers) != len(orientations):
            raise ValueError("Number of identifiers ({}) and orientations ({}) must be the same."
                             .format(len(identifiers), len(orientations)))
        self.identifiers = identifiers
        self.structures = structures
        self.orientations = orientations
        self.struct_lib = dict()
        for ident, struct, ori in zip(identifiers, structures, orientations):
            self.struct_lib[ident] = (struct, ori)

    @classmethod
    def from_orientation_lists(cls, identifiers, structures, orientations):
        return cls(identifiers, structures, orientations)

    def get_library_size(self, to_print=False):
        size = 0
        for i in range(len(self.orientations)):
            n = 1 if len(self.orientations[i]) == 1 else len(self.orientations[i])
            size += n
            if to_print:
                print(self.identifiers[i], n)
        return size


def _entry_from_angles(library, phase, angles):
    for i, orientation in enumerate(library[phase]["orientations"]):
        if np.sum(np.abs(np.subtract(orientation, angles))) < 1e-2:
            return i
    raise ValueError("It appears that no library entry lies with 1e-2 of the target angle")


class DiffractionLibrary(dict):
    """Phase -> {"simulations", "orientations", "pixel_coords", "intensities"}
    (diffsims/libraries/diffraction_library.py:86-178, without pickling)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.identifiers = None
        self.structures = None
        self.diffraction_generator = None
        self.reciprocal_radius = 0.0
        self.with_direct_beam = False

    def get_library_entry(self, phase=None, angle=None):
        if phase is not None:
            entry = self[phase]
            i = _entry_from_angles(self, phase, angle) if angle is not None else 0
        elif angle is not None:
            raise ValueError("To select a certain angle you must first specify a phase")
        else:
            entry = next(iter(self.values()))
            i = 0
        return {"Sim": entry["simulations"][i], "intensities": entry["intensities"][i],
                "pixel_coords": entry["pixel_coords"][i],
                "pattern_norm": np.linalg.norm(entry["intensities"][i])}


class DiffractionLibraryGenerator:
    """Libraries of legacy simulations (library_generator.py:40-153): per
    phase one reflection table and one spot-kernel launch over every
    orientation of its list."""

    def __init__(self, electron_diffraction_calculator):
        self.electron_diffraction_calculator = electron_diffraction_calculator

    def get_diffraction_library(self, structure_library, calibration, reciprocal_radius,
                                half_shape, with_direct_beam=True, max_excitation_error=1e-2,
                                shape_factor_width=None, debye_waller_factors={}):
        lib = DiffractionLibrary()
        gen = self.electron_diffraction_calculator
        if shape_factor_width is None:
            shape_factor_width = max_excitation_error
        for phase_name in structure_library.struct_lib.keys():
            structure, orientations = structure_library.struct_lib[phase_name]
            n = len(orientations)
            tab = gen._table(structure, reciprocal_radius, debye_waller_factors)
            idx, xyz, I, counts = gen._simulate_batch(tab, np.asarray(orientations, dtype=np.float64)
                                                      .reshape(-1, 3), reciprocal_radius,
                                                      max_excitation_error, shape_factor_width)
            hkl_all = tab.hkl_host().astype(np.float64)
            simulations = np.empty(n, dtype="object")
            pixel_coords = np.empty(n, dtype="object")
            intensities = np.empty(n, dtype="object")
            start = 0
            for i, c in enumerate(counts):
                sl = slice(start, start + int(c))
                start += int(c)
                sim = DiffractionSimulation(coordinates=xyz[sl], indices=hkl_all[idx[sl]],
                                            intensities=I[sl], with_direct_beam=with_direct_beam)
                sim.calibration = calibration
                pixel_coords[i] = np.rint(sim.calibrated_coordinates[:, :2] + half_shape).astype(int)
                simulations[i] = sim
                intensities[i] = sim.intensities
            lib[phase_name] = {"simulations": simulations, "orientations": orientations,
                               "pixel_coords": pixel_coords, "intensities": intensities}
        lib.identifiers = structure_library.identifiers
        lib.structures = structure_library.structures
        lib.diffraction_generator = gen
        lib.reciprocal_radius = reciprocal_radius
        lib.with_direct_beam = with_direct_beam
        return lib
