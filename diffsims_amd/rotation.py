"""Rotations (orix.quaternion.Rotation subset) and SO(3) sampling.

Conventions follow orix as the reference uses it:
* ``Rotation.from_euler`` takes Bunge (ZXZ) angles; with the default
  ``direction="lab2crystal"`` the quaternion is Rowenhorst et al. (2015),
  P = +1, with a non-negative scalar part;
* ``to_matrix()`` is the matrix R of v -> q v q^-1, which for those
  quaternions equals the passive Bunge matrix g(phi1, Phi, phi2);
* ``rotation * vectors`` applies R, ``~rotation * vectors`` applies R^T; the
  reference computes lab-frame reciprocal vectors as ``~rotation * miller``
  (diffsims/crystallography/_diffracting_vector.py:160).
"""

from __future__ import annotations

import numpy as np

__all__ = ["Rotation", "uniform_so3_grid"]


def _qmul(p, q):
    w1, x1, y1, z1 = np.moveaxis(p, -1, 0)
    w2, x2, y2, z2 = np.moveaxis(q, -1, 0)
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=-1)


def _matrix(q):
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    qq = w * w - (x * x + y * y + z * z)
    M = np.empty(q.shape[:-1] + (3, 3))
    M[..., 0, 0] = qq + 2 * x * x
    M[..., 0, 1] = 2 * (x * y - w * z)
    M[..., 0, 2] = 2 * (x * z + w * y)
    M[..., 1, 0] = 2 * (y * x + w * z)
    M[..., 1, 1] = qq + 2 * y * y
    M[..., 1, 2] = 2 * (y * z - w * x)
    M[..., 2, 0] = 2 * (z * x - w * y)
    M[..., 2, 1] = 2 * (z * y + w * x)
    M[..., 2, 2] = qq + 2 * z * z
    return M


def _from_matrix(M):
    """Unit quaternions (w >= 0) of proper rotation matrices (Shepperd's method)."""
    M = np.asarray(M, dtype=np.float64)
    flat = M.reshape(-1, 3, 3)
    out = np.empty((flat.shape[0], 4))
    for i, m in enumerate(flat):
        tr = m[0, 0] + m[1, 1] + m[2, 2]
        cand = np.array([tr, m[0, 0], m[1, 1], m[2, 2]])
        k = int(np.argmax(cand))
        if k == 0:
            w = 0.5 * np.sqrt(1 + tr)
            q = [w, (m[2, 1] - m[1, 2]) / (4 * w), (m[0, 2] - m[2, 0]) / (4 * w),
                 (m[1, 0] - m[0, 1]) / (4 * w)]
        elif k == 1:
            x = 0.5 * np.sqrt(1 + m[0, 0] - m[1, 1] - m[2, 2])
            q = [(m[2, 1] - m[1, 2]) / (4 * x), x, (m[0, 1] + m[1, 0]) / (4 * x),
                 (m[0, 2] + m[2, 0]) / (4 * x)]
        elif k == 2:
            y = 0.5 * np.sqrt(1 - m[0, 0] + m[1, 1] - m[2, 2])
            q = [(m[0, 2] - m[2, 0]) / (4 * y), (m[0, 1] + m[1, 0]) / (4 * y), y,
                 (m[1, 2] + m[2, 1]) / (4 * y)]
        else:
            z = 0.5 * np.sqrt(1 - m[0, 0] - m[1, 1] + m[2, 2])
            q = [(m[1, 0] - m[0, 1]) / (4 * z), (m[0, 2] + m[2, 0]) / (4 * z),
                 (m[1, 2] + m[2, 1]) / (4 * z), z]
        q = np.asarray(q)
        out[i] = -q if q[0] < 0 else q
    return out.reshape(M.shape[:-2] + (4,))


class Rotation:
    """Unit-quaternion rotations, shape ``data.shape[:-1]``."""

    def __init__(self, data):
        d = np.asarray(data, dtype=np.float64)
        if d.shape[-1] != 4:
            raise ValueError("quaternion data must have a last axis of length 4")
        self.data = d.reshape((1, 4)) if d.ndim == 1 else d

    # -- construction ------------------------------------------------------
    @classmethod
    def from_euler(cls, euler, direction: str = "lab2crystal", degrees: bool = False):
        e = np.asarray(euler, dtype=np.float64)
        if degrees:
            e = np.deg2rad(e)
        e = e.reshape(-1, 3) if e.ndim == 1 else e
        p1, P, p2 = e[..., 0], e[..., 1], e[..., 2]
        sig = 0.5 * (p1 + p2)
        dlt = 0.5 * (p1 - p2)
        c, s = np.cos(P / 2), np.sin(P / 2)
        q = np.stack([c * np.cos(sig), -s * np.cos(dlt), -s * np.sin(dlt), -c * np.sin(sig)], -1)
        q = np.where(q[..., :1] < 0, -q, q)
        r = cls(q)
        if direction == "crystal2lab":
            r = ~r
        elif direction != "lab2crystal":
            raise ValueError(f"unknown direction {direction!r}")
        return r

    @classmethod
    def from_matrix(cls, matrix):
        return cls(_from_matrix(matrix))

    @classmethod
    def from_axes_angles(cls, axes, angles, degrees: bool = False):
        ax = np.asarray(axes, dtype=np.float64)
        ang = np.asarray(angles, dtype=np.float64)
        if degrees:
            ang = np.deg2rad(ang)
        ax = ax / np.linalg.norm(ax, axis=-1, keepdims=True)
        ang = ang[..., None]
        q = np.concatenate([np.cos(ang / 2) * np.ones(ax.shape[:-1] + (1,)), np.sin(ang / 2) * ax],
                           axis=-1)
        return cls(q.reshape(-1, 4))

    @classmethod
    def identity(cls, shape=(1,)):
        shape = (shape,) if np.isscalar(shape) else tuple(shape)
        q = np.zeros(shape + (4,))
        q[..., 0] = 1.0
        return cls(q)

    @classmethod
    def random(cls, shape=(1,), seed=None):
        """Haar-uniform random rotations (normalised Gaussian 4-vectors)."""
        shape = (shape,) if np.isscalar(shape) else tuple(shape)
        rng = np.random.default_rng(seed)
        q = rng.normal(size=shape + (4,))
        q /= np.linalg.norm(q, axis=-1, keepdims=True)
        q = np.where(q[..., :1] < 0, -q, q)
        return cls(q)

    # -- views ---------------------------------------------------------------
    @property
    def shape(self):
        return self.data.shape[:-1]

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def __len__(self):
        return self.shape[0]

    def flatten(self):
        return Rotation(self.data.reshape(-1, 4))

    def __getitem__(self, key):
        d = self.data[key]
        return Rotation(d.reshape(-1, 4) if d.ndim == 1 else d)

    def __iter__(self):
        flat = self.data.reshape(-1, 4)
        for i in range(flat.shape[0]):
            yield Rotation(flat[i:i + 1])

    def to_matrix(self) -> np.ndarray:
        return _matrix(self.data)

    def to_euler(self, degrees: bool = False) -> np.ndarray:
        """Bunge angles of the passive matrix (inverse of from_euler)."""
        M = self.to_matrix().reshape(-1, 3, 3)
        P = np.arccos(np.clip(M[:, 2, 2], -1, 1))
        p1 = np.arctan2(M[:, 2, 0], -M[:, 2, 1])
        p2 = np.arctan2(M[:, 0, 2], M[:, 1, 2])
        small = np.abs(np.sin(P)) < 1e-12
        p1 = np.where(small, np.arctan2(M[:, 0, 1], M[:, 0, 0]), p1)
        p2 = np.where(small, 0.0, p2)
        e = np.mod(np.stack([p1, P, p2], -1), 2 * np.pi)
        e[:, 1] = P
        e = e.reshape(self.shape + (3,))
        return np.rad2deg(e) if degrees else e

    def __invert__(self):
        return Rotation(self.data * np.array([1.0, -1.0, -1.0, -1.0]))

    def __mul__(self, other):
        if isinstance(other, Rotation):
            return Rotation(_qmul(self.data, other.data))
        v = np.asarray(getattr(other, "data", other), dtype=np.float64)
        M = self.to_matrix()
        if self.size == 1:
            return v @ M.reshape(3, 3).T
        return np.einsum("...ij,...j->...i", M, v)

    def __repr__(self):
        return f"Rotation {self.shape}\n{np.array_str(self.data, precision=4)}"


def uniform_so3_grid(n: int) -> Rotation:
    """Deterministic near-uniform set of n rotations: the super-Fibonacci spiral
    on the unit quaternion sphere (Alexa, CVPR 2022)."""
    phi = np.sqrt(2.0)
    psi = 1.533751168755204288118041
    s = np.arange(n, dtype=np.float64) + 0.5
    t = s / n
    r, R = np.sqrt(t), np.sqrt(1.0 - t)
    a, b = 2 * np.pi * s / phi, 2 * np.pi * s / psi
    q = np.stack([r * np.sin(a), r * np.cos(a), R * np.sin(b), R * np.cos(b)], axis=-1)
    q = np.where(q[..., :1] < 0, -q, q)
    return Rotation(q)
