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

__all__ = ["Rotation", "bunge_from_quaternions", "uniform_so3_grid"]


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


def bunge_from_quaternions(q) -> np.ndarray:
    """Bunge angles (radians) of quaternions [..., 4]: Rowenhorst et al.,
    Modelling Simul. Mater. Sci. Eng. 23 (2015) 083501, eq. A.14 with P = +1,
    with its Phi = 0 and Phi = pi branches; negative phi1 / phi2 are moved
    into [0, 2 pi) and a negative Phi into [0, pi].  Returns [n, 3]."""
    q = np.asarray(q, dtype=np.float64).reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    q03 = w * w + z * z
    q12 = x * x + y * y
    chi = np.sqrt(q03 * q12)
    out = np.zeros((q.shape[0], 3))
    flat = (chi == 0) & (q12 == 0)
    flip = (chi == 0) & (q03 == 0) & ~flat
    gen = ~(flat | flip)
    out[flat, 0] = np.arctan2(-2.0 * w[flat] * z[flat], w[flat] ** 2 - z[flat] ** 2)
    out[flip, 0] = np.arctan2(2.0 * x[flip] * y[flip], x[flip] ** 2 - y[flip] ** 2)
    out[flip, 1] = np.pi
    c = chi[gen]
    W, X, Y, Z = w[gen], x[gen], y[gen], z[gen]
    out[gen, 0] = np.arctan2((X * Z - W * Y) / c, (-W * X - Y * Z) / c)
    out[gen, 1] = np.arctan2(2.0 * c, q03[gen] - q12[gen])
    out[gen, 2] = np.arctan2((W * Y + X * Z) / c, (Y * Z - W * X) / c)
    for k, period in ((0, 2 * np.pi), (1, np.pi), (2, 2 * np.pi)):
        neg = out[:, k] < 0
        out[neg, k] = np.mod(out[neg, k] + 100 * np.pi, period)
    return out


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
        """Bunge angles (inverse of from_euler): Rowenhorst et al. (2015)
        eq. A.14, as orix computes them."""
        e = bunge_from_quaternions(self.data).reshape(self.shape + (3,))
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
