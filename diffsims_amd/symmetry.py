"""Crystallographic point groups for the 1-D profile (host side).

``SimulationGenerator.calculate_diffraction1d`` groups reciprocal lattice
vectors into families of symmetry-equivalent vectors (the reference calls
orix's ``ReciprocalLatticeVector.unique(use_symmetry=True)`` and
``.multiplicity``, diffsims/crystallography/reciprocal_lattice_vector.py:
295-315, 1509-1563).  The 32 point groups are generated here from
generators in the International Tables standard settings, as matrices acting
on Cartesian vectors of the x || a, z || c* frame the reflection tables use:
monoclinic 2-fold / mirror normal along y (unique axis b), trigonal and
hexagonal 3- / 6-fold along z with a 2-fold along x (a), cubic 3-folds along
<111>.  A space-group number maps to its point group by the ranges of the
International Tables.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

__all__ = ["point_group_of_space_group", "point_group_matrices", "symmetry_families"]

# (last space group number of the range, point group)
_SG_RANGES = [(1, "1"), (2, "-1"), (5, "2"), (9, "m"), (15, "2/m"), (24, "222"), (46, "mm2"),
              (74, "mmm"), (80, "4"), (82, "-4"), (88, "4/m"), (98, "422"), (110, "4mm"),
              (122, "-42m"), (142, "4/mmm"), (146, "3"), (148, "-3"), (155, "32"), (161, "3m"),
              (167, "-3m"), (173, "6"), (174, "-6"), (176, "6/m"), (182, "622"), (186, "6mm"),
              (190, "-6m2"), (194, "6/mmm"), (199, "23"), (206, "m-3"), (214, "432"),
              (220, "-43m"), (230, "m-3m")]


def point_group_of_space_group(number: int) -> str:
    """Point group (Hermann-Mauguin) of a space group number 1..230."""
    n = int(number)
    if not 1 <= n <= 230:
        raise ValueError(f"space group number {number} is not in 1..230")
    for last, pg in _SG_RANGES:
        if n <= last:
            return pg
    raise AssertionError


def _rz(n):
    t = 2.0 * np.pi / n
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


_I = np.eye(3)
_INV = -np.eye(3)
_2X = np.diag([1.0, -1.0, -1.0])
_2Y = np.diag([-1.0, 1.0, -1.0])
_2Z = np.diag([-1.0, -1.0, 1.0])
_MX = np.diag([-1.0, 1.0, 1.0])   # mirror normal to x
_MY = np.diag([1.0, -1.0, 1.0])
_MZ = np.diag([1.0, 1.0, -1.0])
_3XYZ = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])  # 3-fold along [111]

_GENERATORS = {
    "1": [_I], "-1": [_INV],
    "2": [_2Y], "m": [_MY], "2/m": [_2Y, _INV],
    "222": [_2Z, _2Y], "mm2": [_2Z, _MX], "mmm": [_2Z, _2Y, _INV],
    "4": [_rz(4)], "-4": [-_rz(4)], "4/m": [_rz(4), _INV], "422": [_rz(4), _2X],
    "4mm": [_rz(4), _MX], "-42m": [-_rz(4), _2X], "4/mmm": [_rz(4), _2X, _INV],
    "3": [_rz(3)], "-3": [_rz(3), _INV], "32": [_rz(3), _2X], "3m": [_rz(3), _MX],
    "-3m": [_rz(3), _2X, _INV],
    "6": [_rz(6)], "-6": [_MZ @ _rz(3)], "6/m": [_rz(6), _INV], "622": [_rz(6), _2X],
    "6mm": [_rz(6), _MX], "-6m2": [_MZ @ _rz(3), _MX], "6/mmm": [_rz(6), _2X, _INV],
    "23": [_2Z, _2X, _3XYZ], "m-3": [_2Z, _2X, _3XYZ, _INV], "432": [_rz(4), _3XYZ],
    "-43m": [-_rz(4), _3XYZ], "m-3m": [_rz(4), _3XYZ, _INV],
}


@lru_cache(maxsize=None)
def _group(name: str) -> tuple:
    if name not in _GENERATORS:
        raise ValueError(f"unknown point group {name!r}")
    elems = [_I]
    keys = {tuple(np.round(_I, 9).ravel())}
    frontier = list(elems)
    while frontier:
        nxt = []
        for a in frontier:
            for g in _GENERATORS[name]:
                m = g @ a
                k = tuple(np.round(m, 9).ravel())
                if k not in keys:
                    keys.add(k)
                    elems.append(m)
                    nxt.append(m)
        frontier = nxt
    return tuple(elems)


def point_group_matrices(name: str) -> np.ndarray:
    """The operations of a point group, (order, 3, 3), on Cartesian vectors."""
    return np.array(_group(str(name)))


def symmetry_families(xyz: np.ndarray, point_group: str):
    """Families of symmetry-equivalent Cartesian vectors.

    Returns (rep, mult): for each family, in the order the reference leaves
    them (descending order of the sorted, 6-decimal orbit), the row of its
    first member in ascending Cartesian order and the number of distinct
    vectors in its orbit."""
    v = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    ops = point_group_matrices(point_group)
    order = np.lexsort((v[:, 2], v[:, 1], v[:, 0]))      # ascending (x, y, z)
    vs = v[order]
    orb = np.einsum("kij,nj->nki", ops, vs).round(6)      # (n, ops, 3)
    # each orbit sorted by z, then y, then x (lexsort's last key first)
    key = np.empty_like(orb)
    for i in range(orb.shape[0]):
        a = orb[i]
        key[i] = a[np.lexsort(a.T)]
    flat = key.reshape(key.shape[0], -1)
    _, first = np.unique(flat, axis=0, return_index=True)
    first = first[::-1]
    mult = np.array([np.unique(orb[i], axis=0).shape[0] for i in first], dtype=np.int64)
    return order[first], mult


# ---- proper rotation groups for SO(3) fundamental zones ----------------------
#
# orix's fundamental zones (get_sample_fundamental, used by the reference's
# get_fundamental_zone_grid) are those of the 11 proper point groups, with
# n-fold axes along z, two-fold axes along x and three-folds along <111>.

_ROTATION_GROUP_GENERATORS = {
    "1": [], "2": [((0, 0, 1), 2)], "222": [((0, 0, 1), 2), ((1, 0, 0), 2)],
    "4": [((0, 0, 1), 4)], "422": [((0, 0, 1), 4), ((1, 0, 0), 2)],
    "3": [((0, 0, 1), 3)], "32": [((0, 0, 1), 3), ((1, 0, 0), 2)],
    "6": [((0, 0, 1), 6)], "622": [((0, 0, 1), 6), ((1, 0, 0), 2)],
    "23": [((0, 0, 1), 2), ((1, 0, 0), 2), ((1, 1, 1), 3)],
    "432": [((0, 0, 1), 4), ((1, 1, 1), 3)],
}

# rotation group of each point group's Laue class
_LAUE_ROTATIONS = {
    "1": "1", "-1": "1", "2": "2", "m": "2", "2/m": "2", "222": "222", "mm2": "222",
    "mmm": "222", "4": "4", "-4": "4", "4/m": "4", "422": "422", "4mm": "422", "-42m": "422",
    "4/mmm": "422", "3": "3", "-3": "3", "32": "32", "3m": "32", "-3m": "32", "6": "6",
    "-6": "6", "6/m": "6", "622": "622", "6mm": "622", "-6m2": "622", "6/mmm": "622",
    "23": "23", "m-3": "23", "432": "432", "-43m": "432", "m-3m": "432",
}


def rotation_group_of_space_group(number: int) -> str:
    """The proper point group whose fundamental zone a space group's
    orientations are sampled in: the rotations of its Laue class."""
    return _LAUE_ROTATIONS[point_group_of_space_group(number)]


def _qmul(p, q):
    a1, b1, c1, d1 = p
    a2, b2, c2, d2 = q
    return np.array([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2])


@lru_cache(maxsize=None)
def _rotation_group(name: str) -> tuple:
    if name not in _ROTATION_GROUP_GENERATORS:
        raise ValueError(f"unknown proper point group {name!r}")
    gens = []
    for axis, n in _ROTATION_GROUP_GENERATORS[name]:
        u = np.asarray(axis, dtype=np.float64)
        u = u / np.linalg.norm(u)
        gens.append(np.concatenate([[np.cos(np.pi / n)], np.sin(np.pi / n) * u]))

    def canon(q):  # q and -q are one rotation
        nz = np.flatnonzero(np.abs(q) > 1e-9)
        return tuple(np.round(q if q[nz[0]] > 0 else -q, 9))

    elems = [np.array([1.0, 0.0, 0.0, 0.0])]
    seen = {canon(elems[0])}
    frontier = list(elems)
    while frontier:
        nxt = []
        for a in frontier:
            for g in gens:
                m = _qmul(g, a)
                k = canon(m)
                if k not in seen:
                    seen.add(k)
                    elems.append(m)
                    nxt.append(m)
        frontier = nxt
    return tuple(elems)


def rotation_group_quaternions(name: str) -> np.ndarray:
    """The rotations of a proper point group as unit quaternions (order, 4),
    the identity first."""
    return np.array(_rotation_group(str(name)))
