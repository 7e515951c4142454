"""Minimal crystal-structure types and the FP64 lattice setup.

The reference consumes ``diffpy.structure`` (Lattice, Atom, Structure) and
``orix.crystal_map.Phase`` objects.  Neither package is available here, so this
module provides small duck-typed equivalents with the attributes the
simulation path reads (diffsims/generators/simulation_generator.py:134-259,
diffsims/utils/sim_utils.py:165-224): ``phase.name``, ``phase.structure``
(iterable of atoms with ``element``, ``xyz``, ``occupancy``),
``phase.structure.lattice`` (``a, b, c, alpha, beta, gamma``, ``rnorm``,
``base``), ``phase.deepcopy()``.  Real diffpy/orix objects work too: the device
path only reads those attributes.

Cartesian frame: x || a, y in the ab-plane, z || c* (the frame orix imposes on
every structure; stated at diffsims/tests/generators/test_simulation_generator.py:311).
"""

from __future__ import annotations

import copy
import math
import re
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

__all__ = ["Lattice", "Atom", "Structure", "Phase", "PhaseArrays", "phase_arrays",
           "reference_frame_positions",
           "lattice_setup", "reciprocal_metric", "highest_hkl", "element_of"]


# ---------------------------------------------------------------------------
# FP64 lattice setup
# ---------------------------------------------------------------------------

_EXACT_COS = {0.0: 1.0, 60.0: 0.5, 90.0: 0.0, 120.0: -0.5, 180.0: -1.0,
              240.0: -0.5, 270.0: 0.0, 300.0: 0.5}


def cosd(x: float) -> float:
    """Cosine of an angle in degrees; exact where the value is 0, +-1/2 or +-1
    (degree-exact trigonometry as diffpy's Lattice uses)."""
    r = math.fmod(float(x), 360.0)
    if r < 0:
        r += 360.0
    v = _EXACT_COS.get(r)
    return v if v is not None else math.cos(math.radians(x))


def sind(x: float) -> float:
    return cosd(90.0 - float(x))


def lattice_setup(abc: Sequence[float], angles: Sequence[float]):
    """(base, rec): direct basis rows and rec = inv(base) in the orix frame.

    The basis is lower triangular in that frame, so its inverse is written out
    in closed form; the columns of ``rec`` are a*, b*, c* and the Cartesian
    reciprocal vector of indices hkl is ``rec @ hkl``.
    """
    a, b, c = (float(v) for v in abc)
    al, be, ga = (float(v) for v in angles)
    ca, cb, cg = cosd(al), cosd(be), cosd(ga)
    sg = sind(ga)
    b1, b2 = b * cg, b * sg
    c1 = c * cb
    c2 = c * (ca - cb * cg) / sg
    c3 = c * math.sqrt(1.0 - cb * cb - ((ca - cb * cg) / sg) ** 2)
    base = np.array([[a, 0.0, 0.0], [b1, b2, 0.0], [c1, c2, c3]])
    rec = np.zeros((3, 3))
    rec[0, 0] = 1.0 / a
    rec[1, 1] = 1.0 / b2
    rec[2, 2] = 1.0 / c3
    rec[1, 0] = -b1 / (a * b2)
    rec[2, 1] = -c2 / (b2 * c3)
    rec[2, 0] = (b1 * c2 - b2 * c1) / (a * b2 * c3)
    return base, rec


def reciprocal_metric(rec: np.ndarray) -> np.ndarray:
    """{G00, G11, G22, G01, G02, G12} of G* = rec^T rec (1/d^2 = hkl G* hkl)."""
    G = rec.T @ rec
    return np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1], G[0, 2], G[1, 2]])


def _dspacing(hkl, m) -> float:
    """1/sqrt(q) with q summed in the order the device enumeration uses."""
    h, k, l = (float(v) for v in hkl)
    q = h * h * m[0]
    q = q + k * k * m[1]
    q = q + l * l * m[2]
    cross = h * k * m[3]
    cross = cross + h * l * m[4]
    cross = cross + k * l * m[5]
    q = q + 2.0 * cross
    return 1.0 / math.sqrt(q)


def highest_hkl(rec: np.ndarray, d_min: float) -> np.ndarray:
    """Per-axis bound of the hkl box: the largest n with d(n e_i) > d_min
    (orix's bound used at diffsims/crystallography/reciprocal_lattice_vector.py:1131;
    the loop of diffsims/crystallography/get_hkl.py:49-57 returns one more)."""
    m = reciprocal_metric(rec)
    out = np.zeros(3, dtype=np.int64)
    for i in range(3):
        e = [0, 0, 0]
        d = d_min + 1.0
        while d > d_min:
            e[i] += 1
            d = _dspacing(e, m)
        out[i] = e[i] - 1
    return out


# ---------------------------------------------------------------------------
# Structure types (diffpy.structure / orix.crystal_map subset)
# ---------------------------------------------------------------------------


class Lattice:
    """Lattice parameters in Angstrom and degrees (diffpy.structure.Lattice subset)."""

    def __init__(self, a=1.0, b=1.0, c=1.0, alpha=90.0, beta=90.0, gamma=90.0):
        self.setLatPar(a, b, c, alpha, beta, gamma)

    def setLatPar(self, a=None, b=None, c=None, alpha=None, beta=None, gamma=None, **_):
        self.a = float(a) if a is not None else self.a
        self.b = float(b) if b is not None else self.b
        self.c = float(c) if c is not None else self.c
        self.alpha = float(alpha) if alpha is not None else self.alpha
        self.beta = float(beta) if beta is not None else self.beta
        self.gamma = float(gamma) if gamma is not None else self.gamma

    def abcABG(self):
        return (self.a, self.b, self.c, self.alpha, self.beta, self.gamma)

    @property
    def base(self) -> np.ndarray:
        return lattice_setup((self.a, self.b, self.c), (self.alpha, self.beta, self.gamma))[0]

    @property
    def recbase(self) -> np.ndarray:
        return lattice_setup((self.a, self.b, self.c), (self.alpha, self.beta, self.gamma))[1]

    def rnorm(self, hkl) -> np.ndarray:
        """|g| of reciprocal-lattice indices (1/Angstrom)."""
        rec = self.recbase
        g = np.asarray(hkl, dtype=np.float64) @ rec.T
        return np.sqrt(np.sum(g * g, axis=-1))

    def __repr__(self):
        return ("Lattice(a={:g}, b={:g}, c={:g}, alpha={:g}, beta={:g}, gamma={:g})"
                .format(*self.abcABG()))


class Atom:
    """An atom site (diffpy.structure.Atom subset): type symbol, fractional xyz,
    occupancy."""

    def __init__(self, atype="", xyz=(0.0, 0.0, 0.0), occupancy=1.0, lattice=None,
                 Uisoequiv=None, Bisoequiv=None, **_):
        self.element = str(atype)
        self.xyz = np.asarray(xyz, dtype=np.float64).reshape(3).copy()
        self.occupancy = float(occupancy)
        self.lattice = lattice
        self.Uisoequiv = Uisoequiv
        self.Bisoequiv = Bisoequiv

    def __repr__(self):
        return f"Atom({self.element!r}, {self.xyz.tolist()}, occupancy={self.occupancy:g})"


class Structure(list):
    """A list of atoms with a lattice (diffpy.structure.Structure subset)."""

    def __init__(self, atoms=None, lattice=None, title=""):
        super().__init__(atoms or [])
        self.lattice = lattice if lattice is not None else Lattice()
        self.title = title

    @property
    def xyz(self) -> np.ndarray:
        return np.array([a.xyz for a in self]).reshape(-1, 3)

    @property
    def element(self):
        return [a.element for a in self]

    @property
    def occupancy(self) -> np.ndarray:
        return np.array([a.occupancy for a in self])

    def copy(self):
        return copy.deepcopy(self)


class Phase:
    """A named crystal phase (orix.crystal_map.Phase subset)."""

    def __init__(self, name="", space_group=None, point_group=None, structure=None, color=None):
        self.name = name
        self.space_group = space_group
        self.point_group = point_group
        self.structure = structure if structure is not None else Structure()

    def deepcopy(self):
        return copy.deepcopy(self)

    def __repr__(self):
        return f"Phase({self.name!r}, {self.structure.lattice!r}, {len(self.structure)} atoms)"


# ---------------------------------------------------------------------------
# Arrays the device path consumes
# ---------------------------------------------------------------------------

_ELEMENT_RE = re.compile(r"^(?P<element>[A-Za-z]*)")


def element_of(symbol: str) -> str:
    """Alphabetic head of a CIF atom-type symbol
    (diffsims/structure_factor/atomic_scattering_parameters.py:145-164)."""
    return _ELEMENT_RE.match(symbol).group("element")


@dataclass
class PhaseArrays:
    """Plain arrays of one phase: lattice parameters and per-atom data."""
    abc: tuple
    angles: tuple
    symbols: list
    xyz: np.ndarray
    occ: np.ndarray
    B: np.ndarray
    name: str = ""


def reference_frame_positions(lattice, xyz: np.ndarray) -> np.ndarray:
    """Fractional atom positions in the frame of the Miller indices.

    orix re-aligns every structure it holds to x || a, z || c* and keeps the
    atoms' Cartesian positions, which rewrites their fractional coordinates;
    the reference undoes that with xyz @ inv(stdbase @ recbase) before the
    structure factor (diffsims/utils/sim_utils.py:288-291).  The same matrix
    is applied here whenever the lattice carries diffpy's ``stdbase`` and
    ``recbase`` (diffpy / orix objects; it is the identity for a lattice that
    was never re-aligned).  This module's own :class:`Lattice` keeps the
    positions as given and has no ``stdbase``, so they pass unchanged.  A
    lattice with a ``baserot`` other than the identity but no ``stdbase``
    cannot be undone and raises ``ValueError``.
    """
    std = getattr(lattice, "stdbase", None)
    rb = getattr(lattice, "recbase", None)
    if std is not None and rb is not None:
        mat = np.linalg.inv(np.asarray(std, dtype=np.float64) @ np.asarray(rb, dtype=np.float64))
        return np.asarray(xyz, dtype=np.float64) @ mat
    br = getattr(lattice, "baserot", None)
    if br is not None and not np.allclose(np.asarray(br, dtype=np.float64), np.eye(3),
                                          rtol=0, atol=1e-12):
        raise ValueError("the lattice has a base rotation but no stdbase/recbase, so the atom "
                         "positions cannot be expressed in the frame of the Miller indices")
    return np.asarray(xyz, dtype=np.float64)


def phase_arrays(phase, debye_waller_factors=None) -> PhaseArrays:
    """Read a phase (this module's, or orix/diffpy) into :class:`PhaseArrays`.

    ``debye_waller_factors`` follows the reference: a dict from element symbol
    to B (A^2), missing elements 0 (diffsims/utils/sim_utils.py:220).  As an
    extension a sequence with one B per atom is accepted (per-atom B cannot be
    expressed through the reference's dict).
    """
    st = phase.structure
    lat = st.lattice
    atoms = list(st)
    symbols = [str(a.element) for a in atoms]
    xyz = reference_frame_positions(
        lat, np.array([np.asarray(a.xyz, dtype=np.float64) for a in atoms]).reshape(-1, 3))
    occ = np.array([float(a.occupancy) for a in atoms], dtype=np.float64)
    if debye_waller_factors is None:
        debye_waller_factors = {}
    if isinstance(debye_waller_factors, dict):
        B = np.array([float(debye_waller_factors.get(element_of(s), 0)) for s in symbols])
    else:
        B = np.asarray(debye_waller_factors, dtype=np.float64).reshape(-1)
        if B.shape[0] != len(atoms):
            raise ValueError(f"{B.shape[0]} per-atom Debye-Waller factors for {len(atoms)} atoms")
    return PhaseArrays((lat.a, lat.b, lat.c), (lat.alpha, lat.beta, lat.gamma), symbols, xyz,
                       occ, B, getattr(phase, "name", ""))
