"""SimulationGenerator: drop-in for diffsims' kinematical 2-D simulation entry.

Same constructor, attributes, errors and ``calculate_diffraction2d`` signature
as diffsims/generators/simulation_generator.py:62-259, returning a
:class:`Simulation2D`.  Per phase the reflection table (hkl box, g, |F|^2) is
built once on the GPU, all orientations are simulated in one persistent
kernel launch, and the spot lists are copied back once.
"""

from __future__ import annotations

import warnings
from typing import Optional

import numpy as np

from . import engine, shape_factor_models as sfm
from .crystallography import phase_arrays
from .diffracting_vector import DiffractingVector
from .rotation import Rotation
from .scattering import electron_wavelength
from .simulation2d import Simulation2D, _is_phase, _is_rotation

__all__ = ["SimulationGenerator"]

_shape_factor_model_mapping = {
    "linear": sfm.linear,
    "atanc": sfm.atanc,
    "sinc": sfm.sinc,
    "sin2c": sfm.sin2c,
    "lorentzian": sfm.lorentzian,
}

# callables with a HIP implementation, by identity
_DEVICE_MODELS = {
    sfm.linear: "linear",
    sfm.atanc: "atanc",
    sfm.sinc: "sinc",
    sfm.sin2c: "sin2c",
    sfm.lorentzian: "lorentzian",
    sfm.binary: "binary",
}


class SimulationGenerator:
    """Kinematic diffraction simulations on an MI355X GPU."""

    def __repr__(self):
        return (f"SimulationGenerator(accelerating_voltage={self.accelerating_voltage}, "
                f"scattering_params={self.scattering_params}, "
                f"approximate_precession={self.approximate_precession})")

    def __init__(self, accelerating_voltage: float = 200, scattering_params: str = "lobato",
                 precession_angle: float = 0, shape_factor_model="lorentzian",
                 approximate_precession: bool = True, minimum_intensity: float = 1e-20,
                 **kwargs):
        self.accelerating_voltage = accelerating_voltage
        self.precession_angle = np.abs(precession_angle)
        self.approximate_precession = approximate_precession
        if isinstance(shape_factor_model, str):
            if shape_factor_model in _shape_factor_model_mapping:
                self.shape_factor_model = _shape_factor_model_mapping[shape_factor_model]
            else:
                raise NotImplementedError(
                    f"{shape_factor_model} is not a recognized shape factor model, choose from: "
                    f"{_shape_factor_model_mapping.keys()} or provide your own function.")
        else:
            self.shape_factor_model = shape_factor_model
        self.minimum_intensity = minimum_intensity
        self.shape_factor_kwargs = kwargs
        if scattering_params in ["lobato", "xtables", None]:
            self.scattering_params = scattering_params
        else:
            raise NotImplementedError(
                "The scattering parameters `{}` is not implemented. See documentation for "
                "available implementations.".format(scattering_params))

    @property
    def wavelength(self) -> float:
        return electron_wavelength(self.accelerating_voltage)

    # -- settings -----------------------------------------------------------------
    def _device_model(self) -> Optional[str]:
        if self.precession_angle != 0 and self.approximate_precession:
            return "lorentzian_precession"
        # built-in models run on the GPU (with precession: their full average)
        return _DEVICE_MODELS.get(self.shape_factor_model)

    def _settings(self, max_excitation_error, shape_factor_width, model=None,
                  minimum_intensity=None) -> engine.SimSettings:
        model = model or self._device_model()
        kw = dict(self.shape_factor_kwargs)
        minima = float(kw.pop("minima_number", 5))
        if kw and model not in (None, "binary"):
            raise TypeError(f"unexpected shape factor arguments {sorted(kw)}")
        return engine.SimSettings(
            accelerating_voltage=self.accelerating_voltage,
            max_excitation_error=max_excitation_error,
            shape_factor_width=shape_factor_width,
            shape_model=model if model != "lorentzian_precession" else "lorentzian",
            minima_number=minima,
            precession_angle=float(self.precession_angle),
            approximate_precession=self.approximate_precession or model == "lorentzian_precession",
            minimum_intensity=self.minimum_intensity if minimum_intensity is None else minimum_intensity)

    # -- the drop-in entry ----------------------------------------------------------
    def calculate_diffraction2d(self, phase, rotation=None, reciprocal_radius: float = 1.0,
                                with_direct_beam: bool = True, max_excitation_error: float = 1e-2,
                                shape_factor_width: float = None, debye_waller_factors: dict = None,
                                show_progressbar: bool = False):
        """Spot lists of one or more phases, each with its own orientations
        (simulation_generator.py:134-259)."""
        if rotation is None:
            rotation = Rotation.from_euler((0, 0, 0), degrees=True)
        if _is_phase(phase):
            phase = [phase]
        if _is_rotation(rotation):
            rotation = [rotation]
        if len(phase) != len(rotation):
            raise ValueError("The number of phases and rotations must be equal. "
                             f"Got {len(phase)} phases and {len(rotation)} rotations.")
        if debye_waller_factors is None:
            debye_waller_factors = {}

        vectors = []
        for p, rot in zip(phase, rotation):
            pbar = None
            if show_progressbar:
                from tqdm import tqdm
                pbar = tqdm(desc=p.name, total=rot.size)
            phase_vectors = self._simulate_phase(p, rot, reciprocal_radius, with_direct_beam,
                                                 max_excitation_error, shape_factor_width,
                                                 debye_waller_factors)
            if pbar is not None:
                pbar.update(rot.size)
                pbar.close()
            vectors.append(phase_vectors)

        if len(phase) == 1:
            vectors = vectors[0]
            phase = phase[0]
            rotation = rotation[0]
            if rotation.size == 1:
                vectors = vectors[0]
        return Simulation2D(phases=phase, coordinates=vectors, rotations=rotation,
                            simulation_generator=self, reciprocal_radius=reciprocal_radius)

    def _simulate_phase(self, p, rot, reciprocal_radius, with_direct_beam, max_excitation_error,
                        shape_factor_width, debye_waller_factors):
        pa = phase_arrays(p, debye_waller_factors)
        tab = engine.build_phase_table(pa, reciprocal_radius, with_direct_beam,
                                       self.scattering_params)
        M = np.asarray(rot.to_matrix(), dtype=np.float64).reshape(-1, 3, 3)
        Rt = engine.rotation_matrices_to_device(M)
        model = self._device_model()
        if model is not None:
            batch = engine.simulate(tab, Rt, self._settings(max_excitation_error, shape_factor_width),
                                    reciprocal_radius)
            return _materialize(batch, p, M)
        return self._simulate_custom(tab, p, M, Rt, reciprocal_radius, max_excitation_error,
                                     shape_factor_width)

    def _simulate_custom(self, tab, p, M, Rt, reciprocal_radius, max_excitation_error,
                         shape_factor_width):
        """Shape models without a HIP kernel (user callables, full precession
        integration): the GPU selects the intersecting reflections and returns
        them uncut; the callable is evaluated on them and the relative cut
        applied (simulation_generator.py:227-240)."""
        settings = self._settings(max_excitation_error, shape_factor_width, model="binary",
                                  minimum_intensity=-1.0)
        batch = engine.simulate(tab, Rt, settings, reciprocal_radius)
        idx, xyz, F2, counts = _flat_spots(batch, M)
        k = 1.0 / self.wavelength
        r = np.sqrt(np.sum(xyz[:, :2] ** 2, axis=1))
        s = (-np.sqrt(k ** 2 - r ** 2) + k) - xyz[:, 2]
        width = shape_factor_width if shape_factor_width is not None else max_excitation_error
        if self.precession_angle == 0:
            sf = self.shape_factor_model(s, width, **self.shape_factor_kwargs)
        else:
            sf = _precession_integral(s, r, np.deg2rad(self.precession_angle),
                                      self.shape_factor_model, width, **self.shape_factor_kwargs)
        I = np.asarray(sf, dtype=np.float64) * F2
        out = []
        start = 0
        hkl_all = tab.hkl_host().astype(np.float64)
        for o, c in enumerate(counts):
            sl = slice(start, start + c)
            start += c
            Io = I[sl]
            if c == 0:
                keep = np.zeros(0, dtype=bool)
            else:
                keep = Io > np.max(Io) * self.minimum_intensity
            dv = DiffractingVector(p, xyz=xyz[sl][keep], intensity=Io[keep], rotation_matrix=M[o])
            dv._hkl = hkl_all[idx[sl][keep]]
            out.append(dv)
        return out

    def get_intersecting_reflections(self, recip, rot, wavelength: float,
                                     max_excitation_error: float, shape_factor_width: float = None,
                                     with_direct_beam: bool = True):
        """Reflections of ``recip`` on the Ewald sphere for one orientation and
        their shape factors (simulation_generator.py:319-412), computed by the
        simulation kernel with |F|^2 = 1 and no cut."""
        import torch

        hkl = np.rint(np.asarray(recip.hkl, dtype=np.float64)).astype(np.int32)
        if with_direct_beam:
            hkl = np.vstack([hkl, np.zeros((1, 3), dtype=np.int32)])
        lat = recip.phase.structure.lattice
        tab = engine.table_from_hkl(hkl, (lat.a, lat.b, lat.c), (lat.alpha, lat.beta, lat.gamma))
        M = np.asarray(rot.to_matrix(), dtype=np.float64).reshape(1, 3, 3)
        Rt = engine.rotation_matrices_to_device(M)
        model = self._device_model()
        if model is None:
            raise NotImplementedError("get_intersecting_reflections needs a shape model with a "
                                      "GPU kernel")
        settings = self._settings(max_excitation_error, shape_factor_width, minimum_intensity=-1.0)
        settings.accelerating_voltage = _voltage_of(wavelength)
        batch = engine.simulate(tab, Rt, settings, max(tab.g_max, 1e-6), cap=max(tab.n, 1))
        idx, xyz, sf, _ = _flat_spots(batch, M)
        vecs = DiffractingVector(recip.phase, xyz=xyz, rotation_matrix=M[0])
        vecs._hkl = hkl[idx].astype(np.float64)
        return vecs, hkl[idx].astype(np.float64), sf

    def calculate_diffraction1d(self, phase, reciprocal_radius: float = 1.0,
                                minimum_intensity: float = 1e-3, debye_waller_factors: dict = None):
        """1-D profile of one phase (simulation_generator.py:261-317): the
        reflections of from_min_dspacing(phase) (d >= 0.7 A, the reference's
        default; ``reciprocal_radius`` only travels to the result) on the
        GPU with their |F|^2, grouped into families of the phase's point
        group (host), intensity |F|^2 x multiplicity scaled to a maximum of
        100, families at or below ``minimum_intensity`` dropped, hkil
        indices for a hexagonal lattice."""
        from .simulation1d import Simulation1D
        from .symmetry import point_group_of_space_group, symmetry_families

        pg = getattr(phase, "point_group", None)
        if pg is not None and not isinstance(pg, str):
            pg = getattr(pg, "name", None)
        if pg is None:
            sg = getattr(phase, "space_group", None)
            if sg is not None and not isinstance(sg, (int, np.integer)):
                sg = getattr(sg, "number", None)
            if sg is None:
                raise ValueError(f"The phase {phase.name!r} must have a point group (or a space "
                                 "group) to reduce its reflections by symmetry")
            pg = point_group_of_space_group(int(sg))
        pa = phase_arrays(phase, debye_waller_factors or {})
        tab = engine.build_phase_table(pa, 1.0 / 0.7, False, self.scattering_params, d_min=0.7)
        hkl = tab.hkl_host().astype(np.int64)
        g = hkl.astype(np.float64) @ tab.rec.T
        rep, mult = symmetry_families(g, pg)
        F2 = tab.F2.double().cpu().numpy()[rep]
        intensities = F2 * mult
        gs = np.sqrt(np.sum(g[rep] ** 2, axis=1))
        hkl_r = hkl[rep].astype(np.float64)
        lat = phase.structure.lattice
        if lat.a == lat.b and lat.alpha == 90 and lat.beta == 90 and lat.gamma == 120:
            hkl_r = np.stack([hkl_r[:, 0], hkl_r[:, 1], -(hkl_r[:, 0] + hkl_r[:, 1]), hkl_r[:, 2]], 1)
        intensities = np.asarray(intensities) / max(intensities) * 100
        mask = intensities > minimum_intensity
        return Simulation1D(phase=phase, reciprocal_spacing=gs[mask],
                            intensities=intensities[mask], hkl=hkl_r[mask],
                            reciprocal_radius=reciprocal_radius, wavelength=self.wavelength)


def _voltage_of(wavelength: float) -> float:
    """Accelerating voltage (kV) whose electron wavelength is ``wavelength``."""
    from .scattering import ELECTRON_MASS as m, ELEMENTARY_CHARGE as e, PLANCK_H as h, \
        SPEED_OF_LIGHT as c
    p2 = (h / (wavelength * 1e-10)) ** 2          # = 2 m e E (1 + e E / (2 m c^2))
    a = e * e / (c * c)
    b = 2 * m * e
    E = (-b + np.sqrt(b * b + 4 * a * p2)) / (2 * a)
    return float(E / 1e3)


def _flat_spots(batch, M):
    """Device spot lists -> flat host arrays (index, FP64 lab xyz, I, counts).
    Lab coordinates are recomputed in FP64 from the integer hkl and the FP64
    orientation matrices."""
    import torch

    n, cap = batch.n, batch.cap
    counts = batch.count.cpu().numpy().astype(np.int64)
    if n == 0 or counts.sum() == 0:
        return (np.zeros(0, np.int64), np.zeros((0, 3)), np.zeros(0), counts)
    mask = torch.arange(cap, device=batch.count.device)[None, :] < batch.count[:, None]
    idx = batch.idx[mask].cpu().numpy().astype(np.int64)
    I = batch.spots[..., 3][mask].cpu().numpy().astype(np.float64)
    hkl = batch.table.hkl_host().astype(np.float64)[idx]
    g = hkl @ batch.table.rec.T
    orient = np.repeat(np.arange(n), counts)
    xyz = np.einsum("ni,nij->nj", g, M[orient])
    return idx, xyz, I, counts


def _materialize(batch, phase, M):
    """One DiffractingVector per orientation (reference result objects)."""
    idx, xyz, I, counts = _flat_spots(batch, M)
    hkl_all = batch.table.hkl_host().astype(np.float64)
    out = []
    start = 0
    for o, c in enumerate(counts):
        sl = slice(start, start + int(c))
        start += int(c)
        dv = DiffractingVector(phase, xyz=xyz[sl], intensity=I[sl], rotation_matrix=M[o])
        dv._hkl = hkl_all[idx[sl]]
        out.append(dv)
    return out


def _precession_integral(s, r, phi, model, width, **kw):
    """(1/2pi) int_0^2pi model(s + r phi cos t) dt per spot with scipy's
    adaptive quadrature (shape_factor_models.py:222-268), for callables that
    have no HIP kernel."""
    from scipy.integrate import quad

    out = np.zeros(np.shape(s))
    for i, (si, ri) in enumerate(zip(np.asarray(s), np.asarray(r))):
        out[i] = quad(lambda t: model(si + ri * phi * np.cos(t), width, **kw), 0, 2 * np.pi)[0] / (2 * np.pi)
    return out
