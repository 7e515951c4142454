"""DiffractingVector: reciprocal-lattice vectors of one simulated pattern.

Host-side result type with the surface of
diffsims/crystallography/_diffracting_vector.py:24-194 that the simulation
path and its consumers use: ``data`` (lab-frame Cartesian, 1/A), ``hkl``,
``intensity``, ``gspacing``, ``size``, slicing, ``to_flat_polar`` and
``rotate_with_basis``.
"""

from __future__ import annotations

import copy

import numpy as np

__all__ = ["DiffractingVector"]


class DiffractingVector:
    """Vectors in the lab frame of a (possibly rotated) phase.

    ``rotation_matrix`` is the orix ``to_matrix()`` M of the orientation the
    vectors were rotated by (lab = M^T crystal); ``hkl`` maps back through it.
    """

    def __init__(self, phase, xyz=None, hkl=None, intensity=None, rotation_matrix=None):
        self.phase = phase
        self.rotation_matrix = (np.eye(3) if rotation_matrix is None
                                else np.asarray(rotation_matrix, dtype=np.float64).reshape(3, 3))
        if (xyz is None) == (hkl is None):
            raise ValueError("Exactly one of `xyz` or `hkl` must be passed")
        if xyz is not None:
            data = np.asarray(xyz)
            self._data = data.reshape(-1, 3) if data.ndim != 2 else data
            self._hkl = None
        else:
            h = np.asarray(hkl, dtype=np.float64).reshape(-1, 3)
            rec = phase.structure.lattice.recbase
            self._data = (h @ rec.T) @ self.rotation_matrix
            self._hkl = h
        if intensity is None:
            self._intensity = np.full(self.size, np.nan)
        else:
            intensity = np.asarray(intensity, dtype=np.float64).reshape(-1)
            if intensity.size != self.size:
                raise ValueError("Length of intensity array must match number of vectors")
            self._intensity = intensity

    # -- array views -----------------------------------------------------------
    @property
    def data(self) -> np.ndarray:
        return self._data

    @data.setter
    def data(self, value):
        # in-place like orix: keeps the dtype (integer coordinates stay integer)
        self._data[...] = value

    def __setitem__(self, key, value):
        self._data[key] = value

    @property
    def size(self) -> int:
        return int(self._data.shape[0])

    @property
    def shape(self):
        return (self.size,)

    def __len__(self):
        return self.size

    @property
    def x(self):
        return self._data[:, 0]

    @property
    def y(self):
        return self._data[:, 1]

    @property
    def z(self):
        return self._data[:, 2]

    @property
    def hkl(self) -> np.ndarray:
        """Miller indices (float), from the stored indices or mapped back through
        the rotation and the lattice."""
        if self._hkl is not None:
            return self._hkl
        base = self.phase.structure.lattice.base
        return (np.asarray(self._data, dtype=np.float64) @ self.rotation_matrix.T) @ base.T

    @property
    def gspacing(self) -> np.ndarray:
        return np.sqrt(np.sum(np.asarray(self._data, dtype=np.float64) ** 2, axis=1))

    @property
    def dspacing(self) -> np.ndarray:
        with np.errstate(divide="ignore"):
            return 1.0 / self.gspacing

    @property
    def intensity(self) -> np.ndarray:
        return self._intensity

    @intensity.setter
    def intensity(self, value):
        if not hasattr(value, "__len__"):
            value = np.array([value] * self.size)
        if len(value) != self.size:
            raise ValueError("Length of intensity array must match number of vectors")
        self._intensity = np.array(value)

    # -- operations --------------------------------------------------------------
    def __getitem__(self, key):
        new = copy.copy(self)
        new._data = np.atleast_2d(self._data[key])
        new._hkl = None if self._hkl is None else np.atleast_2d(self._hkl[key])
        inten = self._intensity[key]
        new._intensity = np.atleast_1d(np.asarray(inten))
        return new

    def flatten(self):
        return self

    def deepcopy(self):
        return copy.deepcopy(self)

    def to_flat_polar(self):
        """(r, theta) of the vectors projected on the x-y plane
        (_diffracting_vector.py:186-194)."""
        d = np.asarray(self._data, dtype=np.float64)
        return np.linalg.norm(d[:, :2], axis=1), np.arctan2(d[:, 1], d[:, 0])

    def rotate_with_basis(self, rotation):
        """Rotate the vectors (lab = M^T v) and compose the basis rotation
        (_diffracting_vector.py:127-161)."""
        if rotation.size != 1:
            raise ValueError("Rotation must be a single rotation")
        M = np.asarray(rotation.to_matrix(), dtype=np.float64).reshape(3, 3)
        new = copy.copy(self)
        new._data = np.asarray(self._data, dtype=np.float64) @ M
        new.rotation_matrix = self.rotation_matrix @ M
        new._intensity = self._intensity.copy()
        return new

    def __repr__(self):
        name = getattr(self.phase, "name", "")
        return f"DiffractingVector {self.shape}, {name}\n" + np.array_str(
            np.asarray(self._data, dtype=np.float64), precision=4, suppress_small=True)
